// HIP kernels of liblfric_hip for gfx950 (MI355X): synthetic operator generation, operator
// layout changes, the colour-ordered LMA apply and its siblings, vector updates and
// deterministic single-pass reductions.
//
// Device operator layout ("level-aligned AoSoA").  Cells are taken in colour-major order
// (pos); column pos owns the elements ebase[pos] + m, m = 0..nlayers (one slot per vertex
// level), colours start on 64-element boundaries.  Entry (df1,df2) of 3D cell (pos, k) lives
// at element e = ebase[pos] + k + top(df1), where top(df1) = 1 when df1 is the top dof of a
// column pair (so both rows that feed output level m sit at the same e).  With p = df1*ndf2 +
// df2 and V = 16 bytes / sizeof(entry), the entry sits at (e/64)*gs + ((p/V)*64 + e%64)*V +
// p%V, gs = 64 * (ndf1*ndf2 rounded up to V).  One wavefront of (cell, level) threads thus
// reads each 16-byte chunk of its 64 elements as one contiguous 1 KiB block, and all chunks of
// a wavefront form one contiguous block.
//
// LMA apply thread decomposition.  The reference kernel adds A_k x_k column by column into
// lhs(map1(df)+k) (components/science/source/kernel/algebra/matrix_vector_kernel_mod.F90:
// 130-138); a top entity of layer k is the bottom entity of layer k+1 (map(top) = map(bottom)
// + 1).  Here one thread owns the vertex level m in [0, nlayers] of one cell column and
// computes, for every bottom-role dof, row(df) of layer m plus row(top partner) of layer m-1.
// Each output dof of a column is therefore written by exactly one thread of the cell, and the
// colouring (no two cells of one colour share an entity) makes one colour pass race-free
// without atomics.  The first colour that reaches a column writes it (no zero fill), later
// colours add to it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "device_common.h"
#include "kernels.h"
#include "colour_apply.h"
#include "mfma_apply.h"
#include "rng.h"
#include "sym27_apply.h"

namespace lfric {

namespace {

// ---------------------------------------------------------------------------------------
// synthetic operators
// ---------------------------------------------------------------------------------------
#pragma clang fp contract(off)
// A_e = B_e B_e^T + shift*I, one wavefront per element (cell position, layer); B_e staged in
// LDS.  Identical arithmetic to the CPU definition: B(i,j) = normal*0.125*sc(j),
// sc(j) = exp((4u-2)*ln10) when scaled, A(i,j) = fma chain over m in order, + shift on the
// diagonal, mirrored.
template <typename T, int NDF>
__global__ __launch_bounds__(kBlock) void k_gen_spd(T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                    const int *__restrict__ cell_of_pos, int nl, long long nelem,
                                                    unsigned long long seed, int scaled, double shift, uint32_t row_top) {
  constexpr int kWaves = kBlock / 64;
  __shared__ double B[kWaves][NDF * NDF];
  __shared__ double sc[kWaves][NDF];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long e = static_cast<long long>(blockIdx.x) * kWaves + w;
  const bool active = e < nelem;
  long long ic = 0, el = 0;
  if (active) {
    const long long pos = e / nl;
    const int k = static_cast<int>(e - pos * nl);
    ic = static_cast<long long>(cell_of_pos[pos]) * nl + k;  // reference 3D cell index
    el = ebase[pos] + k;
  }
  const double ln10 = 0x1.26bb1bbb55516p+1;
  if (active) {
    for (int j = lane; j < NDF; j += 64)
      sc[w][j] = scaled ? rng::exp_small((4.0 * rng::uniform(seed, rng::kStreamScale,
                                                               static_cast<uint64_t>(ic) * NDF + j) - 2.0) * ln10)
                        : 1.0;
  }
  __syncthreads();
  if (active) {
    for (int q = lane; q < NDF * NDF; q += 64) {
      const int j = q % NDF;
      B[w][q] = rng::normal(seed, rng::kStreamB, static_cast<uint64_t>(ic) * NDF * NDF + q) * 0.125 * sc[w][j];
    }
  }
  __syncthreads();
  if (active) {
    for (int q = lane; q < NDF * NDF; q += 64) {
      const int i = q / NDF, j = q % NDF;
      if (j < i) continue;
      double acc = 0.0;
      for (int m = 0; m < NDF; ++m) acc = fma(B[w][i * NDF + m], B[w][j * NDF + m], acc);
      if (i == j) acc += shift;
      A[aidx<T>(gs, el + ((row_top >> i) & 1u), i * NDF + j)] = static_cast<T>(acc);
      A[aidx<T>(gs, el + ((row_top >> j) & 1u), j * NDF + i)] = static_cast<T>(acc);
    }
  }
}

// Entry (df1,df2) of element (pos,k) = sign(df2) * (lo + (hi-lo)*u), counter = ic*n1*n2 + df1*n2 + df2
template <typename T>
__global__ void k_gen_uniform(T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                              const int *__restrict__ cell_of_pos, int nl, long long nelem, int n1, int n2,
                              unsigned long long seed, unsigned long long stream, double lo, double hi,
                              const double *__restrict__ sign, uint32_t row_top) {
  const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const long long pos = e / nl;
  const int k = static_cast<int>(e - pos * nl);
  const long long ic = static_cast<long long>(cell_of_pos[pos]) * nl + k;
  const long long el = ebase[pos] + k;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n2; ++j) {
      const double u = rng::uniform(seed, stream, static_cast<uint64_t>(ic) * n1 * n2 + i * n2 + j);
      double v = lo + (hi - lo) * u;
      if (sign) v = v * sign[j];
      A[aidx<T>(gs, el + ((row_top >> i) & 1u), i * n2 + j)] = static_cast<T>(v);
    }
}

template <typename T>
__global__ void k_vec_uniform(T *__restrict__ v, long long n, unsigned long long seed, unsigned long long stream) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    v[i] = static_cast<T>(2.0 * rng::uniform(seed, stream, static_cast<uint64_t>(i)) - 1.0);
}
// Per-cell dense inverse of a square operator: invert_local_operator_code
// (components/science/source/kernel/algebra/sci_invert_local_operator_kernel_mod.F90:51-77)
// with matrix_invert_lu (infrastructure/source/utilities/matrix_invert_mod.F90:62-216):
// Doolittle LU without pivoting (l below the diagonal, u on and above, in one array), then
// column m of the inverse by forward substitution on e_m and back substitution, in the
// reference's loop order and without contraction, so each entry is the reference's to the
// bit.  One thread per 3D cell.  Up to 8 dofs the factors live in registers; larger elements
// keep them in a workspace laid out [entry][element] so that a wavefront's accesses coalesce.
template <typename T, int N>
struct InvRegs {
  T c[N][N], y[N], x[N];
  __device__ T &C(int i, int j) { return c[i][j]; }
  __device__ T &Y(int i) { return y[i]; }
  __device__ T &X(int i) { return x[i]; }
};
template <typename T>
struct InvWork {
  T *w;
  long long stride, e;
  int n;
  __device__ T &C(int i, int j) { return w[static_cast<size_t>(i * n + j) * stride + e]; }
  __device__ T &Y(int i) { return w[static_cast<size_t>(n * n + i) * stride + e]; }
  __device__ T &X(int i) { return w[static_cast<size_t>(n * n + n + i) * stride + e]; }
};

template <typename T, class S>
__device__ inline void invert_lu(S &s, int n) {
  for (int m = 0; m < n - 1; ++m)
    for (int i = m + 1; i < n; ++i) {
      const T coeff = -(s.C(i, m) / s.C(m, m));
      s.C(i, m) = -coeff;  // l(i,m)
      for (int j = m + 1; j < n; ++j) s.C(i, j) = s.C(i, j) + coeff * s.C(m, j);
    }
}

template <typename T, class S, class Out>
__device__ inline void invert_columns(S &s, int n, Out &&out) {
  for (int m = 0; m < n; ++m) {
    s.Y(0) = m == 0 ? T(1) : T(0);
    for (int i = 1; i < n; ++i) {
      s.Y(i) = i == m ? T(1) : T(0);
      for (int j = 0; j < i; ++j) s.Y(i) = s.Y(i) - s.Y(j) * s.C(i, j);
    }
    s.X(n - 1) = s.Y(n - 1) / s.C(n - 1, n - 1);
    for (int i = n - 2; i >= 0; --i) {
      s.X(i) = s.Y(i);
      for (int j = n - 1; j >= i + 1; --j) s.X(i) = s.X(i) - s.X(j) * s.C(i, j);
      s.X(i) = s.X(i) / s.C(i, i);
    }
    for (int i = 0; i < n; ++i) out(i, m, s.X(i));
  }
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_invert_local(const T *__restrict__ A, T *__restrict__ B, size_t gs,
                                                         const int *__restrict__ ebase, int nl, long long nelem,
                                                         uint32_t row_top) {
  const long long e = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= nelem) return;
  const long long pos = e / nl;
  const long long el = ebase[pos] + (e - pos * nl);
  InvRegs<T, N> s;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) s.c[i][j] = A[aidx<T>(gs, el + ((row_top >> i) & 1u), i * N + j)];
  invert_lu<T>(s, N);
  invert_columns<T>(s, N, [&](int i, int m, T v) { B[aidx<T>(gs, el + ((row_top >> i) & 1u), i * N + m)] = v; });
}

// elements [e0, e0 + count) with the factors in a workspace of (n*n + 2n) x count entries
template <typename T>
__global__ __launch_bounds__(kBlock) void k_invert_local_ws(const T *__restrict__ A, T *__restrict__ B, size_t gs,
                                                            const int *__restrict__ ebase, int nl, long long e0,
                                                            long long count, int n, uint32_t row_top, T *work) {
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= count) return;
  const long long e = e0 + t;
  const long long pos = e / nl;
  const long long el = ebase[pos] + (e - pos * nl);
  InvWork<T> s{work, count, t, n};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) s.C(i, j) = A[aidx<T>(gs, el + ((row_top >> i) & 1u), i * n + j)];
  invert_lu<T>(s, n);
  invert_columns<T>(s, n, [&](int i, int m, T v) { B[aidx<T>(gs, el + ((row_top >> i) & 1u), i * n + m)] = v; });
}

// Lowest-order W0 mass matrix of every cell from Cartesian vertex coordinates:
// compute_mass_matrix_w_scalar_code (components/science/source/kernel/fem/
// sci_compute_mass_matrix_kernel_w_scalar_mod.F90:290-384) with coordinate_jacobian for
// coord_system_xyz (components/science/source/kernel/geometry/sci_coordinate_jacobian_mod.F90:
// 289-445).  One thread per 3D cell: the Jacobian jac(i,dir) = sum_df chi_i(df) d_dir N_df and
// its cofactor determinant at each quadrature point, then the entries df >= df2 as sums over
// (vertical point outer, horizontal point inner) of wh*wv*N_df*N_df2*dj, mirrored.  The basis
// values and weights are tables built on the host ([8][nh*nv] values, [3][8][nh*nv] derivatives,
// [nh] and [nv] weights) and staged in LDS; no contraction, so the arithmetic is the reference's.
template <int NQ>
__global__ __launch_bounds__(kBlock) void k_assemble_mass_w0(double *__restrict__ A, size_t gs,
                                                             const int *__restrict__ ebase,
                                                             const int *__restrict__ map,
                                                             const double *__restrict__ chi1,
                                                             const double *__restrict__ chi2,
                                                             const double *__restrict__ chi3,
                                                             const double *__restrict__ tables, int nl,
                                                             long long nelem, uint32_t row_top) {
  constexpr int NH = NQ * NQ, NP = NH * NQ;
  constexpr int TB = 8 * NP + 24 * NP + NH + NQ;  // values, derivatives, weights
  __shared__ double tab[TB];
  for (int i = threadIdx.x; i < TB; i += kBlock) tab[i] = tables[i];
  __syncthreads();
  const double *b = tab, *db = tab + 8 * NP, *wh = tab + 32 * NP, *wv = wh + NH;
  const long long e = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= nelem) return;
  const long long pos = e / nl;
  const int k = static_cast<int>(e - pos * nl);
  const long long el = ebase[pos] + k;
  double c[3][8];
#pragma unroll
  for (int df = 0; df < 8; ++df) {
    const size_t d = static_cast<size_t>(map[pos * 8 + df]) + k;
    c[0][df] = chi1[d];
    c[1][df] = chi2[d];
    c[2][df] = chi3[d];
  }
  // points in the reference's order (vertical outer, horizontal inner); every entry's sum runs
  // over them in that order, so interleaving the 36 sums leaves each one the reference's
  double s[36];
#pragma unroll
  for (int t = 0; t < 36; ++t) s[t] = 0.0;
  for (int kz = 0; kz < NQ; ++kz)
    for (int ic = 0; ic < NH; ++ic) {
      const int q = kz * NH + ic;
      double J[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
      for (int df = 0; df < 8; ++df)
#pragma unroll
        for (int dir = 0; dir < 3; ++dir)
#pragma unroll
          for (int r = 0; r < 3; ++r) J[r][dir] = J[r][dir] + c[r][df] * db[(dir * 8 + df) * NP + q];
      const double dj = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      const double w = wh[ic] * wv[kz];
      double bq[8];
#pragma unroll
      for (int df = 0; df < 8; ++df) bq[df] = b[df * NP + q];
      int t = 0;
#pragma unroll
      for (int df2 = 0; df2 < 8; ++df2)
#pragma unroll
        for (int df = df2; df < 8; ++df, ++t) s[t] = s[t] + w * bq[df] * bq[df2] * dj;
    }
  int t = 0;
#pragma unroll
  for (int df2 = 0; df2 < 8; ++df2)
#pragma unroll
    for (int df = df2; df < 8; ++df, ++t) {
      A[aidx<double>(gs, el + ((row_top >> df) & 1u), df * 8 + df2)] = s[t];
      A[aidx<double>(gs, el + ((row_top >> df2) & 1u), df2 * 8 + df)] = s[t];
    }
}

// Lowest-order W2 dofs W, S, E, N, B, T: the one nonzero component of each basis vector, and
// its divergence (basis_setup of function_space_constructor_helper_functions_mod.F90:973-1070
// with the face normals +i, -j, +i, -j, +k, +k of reference_element_mod.F90:1436-1441).
__device__ constexpr int kW2Comp[6] = {0, 1, 0, 1, 2, 2};
__device__ constexpr double kW2Div[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};

enum FemKind { kFemMassW3 = 0, kFemMassW2 = 1, kFemDiv = 2, kFemMassWtheta = 3 };

// W3 and W2 mass matrices and the weak divergence W2 -> W3 of every cell from Cartesian
// coordinate fields on a W0 space of the same mesh (map: that space's [pos][8] column-base
// dofs):
//   kFemMassW3  compute_mass_matrix_w3_code (components/science/source/kernel/fem/
//               sci_compute_mass_matrix_kernel_w3_mod.F90:270-356): sum of wh*wv*1*1*dj
//               (rehabilitated) or wh*wv*1*1/dj;
//   kFemMassW2  compute_mass_matrix_w2_code (sci_compute_mass_matrix_kernel_w2_mod.F90:83-164):
//               entries df >= df2 summed from wh*wv*dot_product(J b_df, J b_df2)/dj, mirrored;
//               J b is v*J(:, c) for a basis vector whose one nonzero component v sits in c;
//   kFemDiv     compute_div_operator_code (sci_compute_div_operator_kernel_mod.F90:120-150):
//               sums of wh*wv*1*div(b_df2), over dj without rehabilitation;
//   kFemMassWtheta  compute_mass_matrix_w_scalar_code (sci_compute_mass_matrix_kernel_w_scalar_mod.F90:
//               290-384) with the Wtheta basis 1 - z, z (the W2 B and T factors): entries df >= df2
//               summed from wh*wv*b_df*b_df2*dj, mirrored.
// Quadrature points in the reference's order (vertical outer, horizontal inner), one thread per
// 3D cell, no contraction: each entry is the reference's sum in the reference's order.  Tables:
// coordinate-basis derivatives [3][8][NP], weights [NH] and [NQ], W2 values [6][NP].
template <int NQ, int KIND>
__global__ __launch_bounds__(kBlock) void k_assemble_fem(double *__restrict__ A, size_t gs,
                                                         const int *__restrict__ ebase,
                                                         const int *__restrict__ map,
                                                         const double *__restrict__ chi1,
                                                         const double *__restrict__ chi2,
                                                         const double *__restrict__ chi3,
                                                         const double *__restrict__ tables, int nl,
                                                         long long nelem, uint32_t row_top, int rehab) {
  constexpr int NH = NQ * NQ, NP = NH * NQ;
  constexpr int TB = 24 * NP + NH + NQ + 6 * NP;
  __shared__ double tab[TB];
  for (int i = threadIdx.x; i < TB; i += kBlock) tab[i] = tables[i];
  __syncthreads();
  const double *db = tab, *wh = tab + 24 * NP, *wv = wh + NH, *w2 = wv + NQ;
  const long long e = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (e >= nelem) return;
  const long long pos = e / nl;
  const int k = static_cast<int>(e - pos * nl);
  const long long el = ebase[pos] + k;
  double c[3][8];
#pragma unroll
  for (int df = 0; df < 8; ++df) {
    const size_t d = static_cast<size_t>(map[pos * 8 + df]) + k;
    c[0][df] = chi1[d];
    c[1][df] = chi2[d];
    c[2][df] = chi3[d];
  }
  constexpr int NS = KIND == kFemMassW3 ? 1 : KIND == kFemMassW2 ? 21 : KIND == kFemMassWtheta ? 3 : 6;
  double s[NS];
#pragma unroll
  for (int t = 0; t < NS; ++t) s[t] = 0.0;
  for (int kz = 0; kz < NQ; ++kz)
    for (int ic = 0; ic < NH; ++ic) {
      const int q = kz * NH + ic;
      double J[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
      for (int df = 0; df < 8; ++df)
#pragma unroll
        for (int dir = 0; dir < 3; ++dir)
#pragma unroll
          for (int r = 0; r < 3; ++r) J[r][dir] = J[r][dir] + c[r][df] * db[(dir * 8 + df) * NP + q];
      const double dj = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      const double w = wh[ic] * wv[kz];
      if constexpr (KIND == kFemMassW3) {
        s[0] = s[0] + (rehab ? w * dj : w / dj);
      } else if constexpr (KIND == kFemMassWtheta) {
        const double bq[2] = {w2[4 * NP + q], w2[5 * NP + q]};
        s[0] = s[0] + w * bq[0] * bq[0] * dj;
        s[1] = s[1] + w * bq[1] * bq[0] * dj;
        s[2] = s[2] + w * bq[1] * bq[1] * dj;
      } else if constexpr (KIND == kFemDiv) {
#pragma unroll
        for (int d = 0; d < 6; ++d) {
          const double v = w * kW2Div[d];
          s[d] = s[d] + (rehab ? v : v / dj);
        }
      } else {
        double jb[6][3];
#pragma unroll
        for (int d = 0; d < 6; ++d) {
          const double v = w2[d * NP + q];
#pragma unroll
          for (int r = 0; r < 3; ++r) jb[d][r] = J[r][kW2Comp[d]] * v;
        }
        int t = 0;
#pragma unroll
        for (int df2 = 0; df2 < 6; ++df2)
#pragma unroll
          for (int df = df2; df < 6; ++df, ++t) {
            const double dot = jb[df][0] * jb[df2][0] + jb[df][1] * jb[df2][1] + jb[df][2] * jb[df2][2];
            s[t] = s[t] + w * dot / dj;
          }
      }
    }
  if constexpr (KIND == kFemMassW3) {
    A[aidx<double>(gs, el, 0)] = s[0];
  } else if constexpr (KIND == kFemMassWtheta) {
    const long long e0 = el + (row_top & 1u), e1 = el + ((row_top >> 1) & 1u);
    A[aidx<double>(gs, e0, 0)] = s[0];
    A[aidx<double>(gs, e0, 1)] = s[1];
    A[aidx<double>(gs, e1, 2)] = s[1];
    A[aidx<double>(gs, e1, 3)] = s[2];
  } else if constexpr (KIND == kFemDiv) {
#pragma unroll
    for (int d = 0; d < 6; ++d) A[aidx<double>(gs, el, d)] = s[d];
  } else {
    int t = 0;
#pragma unroll
    for (int df2 = 0; df2 < 6; ++df2)
#pragma unroll
      for (int df = df2; df < 6; ++df, ++t) {
        A[aidx<double>(gs, el + ((row_top >> df) & 1u), df * 6 + df2)] = s[t];
        A[aidx<double>(gs, el + ((row_top >> df2) & 1u), df2 * 6 + df)] = s[t];
      }
  }
}

#pragma clang fp contract(on)

// reference layout matrix(ncell_3d, ndf1, ndf2) <-> device layout
template <typename T, typename S>
__global__ void k_from_reference(T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                 const S *__restrict__ R, const int *__restrict__ cell_of_pos, int nl,
                                 long long nelem, int n1, int n2, uint32_t row_top) {
  const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const long long pos = e / nl;
  const int k = static_cast<int>(e - pos * nl);
  const long long ic = static_cast<long long>(cell_of_pos[pos]) * nl + k;
  const long long el = ebase[pos] + k;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n2; ++j)
      A[aidx<T>(gs, el + ((row_top >> i) & 1u), i * n2 + j)] =
          static_cast<T>(R[ic + nelem * (i + static_cast<long long>(n1) * j)]);
}

template <typename T, typename D>
__global__ void k_to_reference(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                               D *__restrict__ R, const int *__restrict__ cell_of_pos, int nl, long long nelem,
                               int n1, int n2, uint32_t row_top) {
  const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  const long long pos = e / nl;
  const int k = static_cast<int>(e - pos * nl);
  const long long ic = static_cast<long long>(cell_of_pos[pos]) * nl + k;
  const long long el = ebase[pos] + k;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n2; ++j)
      R[ic + nelem * (i + static_cast<long long>(n1) * j)] =
          static_cast<D>(to_d(A[aidx<T>(gs, el + ((row_top >> i) & 1u), i * n2 + j)]));
}

// operator precision change; the chunk width differs between precisions, so entries move
// between layouts (padding entries of dst stay as they are: zero from creation)
template <typename TD, typename TS>
__global__ void k_op_convert(TD *__restrict__ d, size_t gsd, const TS *__restrict__ s, size_t gss, long long nelem,
                             int np) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nelem * np) return;
  const int p = static_cast<int>(t / nelem);
  const long long e = t - static_cast<long long>(p) * nelem;
  d[aidx<TD>(gsd, e, p)] = static_cast<TD>(s[aidx<TS>(gss, e, p)]);
}

template <typename TD, typename TS>
__global__ void k_convert(TD *__restrict__ d, const TS *__restrict__ s, long long n, double scale) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    d[i] = static_cast<TD>(scale == 1.0 ? to_d(s[i]) : to_d(s[i]) * scale);
}

// ---------------------------------------------------------------------------------------
// LMA apply, one colour per launch (see file header)
// ---------------------------------------------------------------------------------------
// Column roles of the output space.  Compile-time tables for the spaces the library builds
// (the dofmap entity order vol, faces W,S,E,N,B,T, edges WB..NB, SW..NW, WT..NT, vertices
// SWB..NWB, SWT..NWT), so every element-matrix load of a thread can be issued before the
// first FMA; RolesRuntime carries any other pattern found in a caller's dofmap.
struct RolesRuntime {
  static constexpr bool kStatic = false;
  Roles r;
  __device__ bool top(int i) const { return (r.top_mask >> i) & 1u; }
  __device__ int partner(int i) const { return r.partner[i]; }
};

template <typename T, int N1, int N2, class R>
__global__ __launch_bounds__(kBlock) void k_lma_colour(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                       const T *__restrict__ x, T *__restrict__ y,
                                                       const int *__restrict__ map1, const int *__restrict__ map2,
                                                       const uint32_t *__restrict__ first, int pos0, int npos, int nl,
                                                       R roles, const int *__restrict__ skip, int xwin) {
  if (skip && *skip) return;
  const int H = nl + 1;
  const int t = xcd_block(blockIdx.x, gridDim.x, xwin) * kBlock + threadIdx.x;
  if (t >= npos * H) return;
  const int lp = t / H;
  const int m = t - lp * H;
  const int pos = pos0 + lp;
  const int *m1 = map1 + static_cast<size_t>(pos) * N1;
  const int *m2 = map2 + static_cast<size_t>(pos) * N2;
  const bool has_m = m < nl, has_p = m > 0;
  const long long e = static_cast<long long>(ebase[pos]) + m;
  T xm[N2], xp[N2];
#pragma unroll
  for (int j = 0; j < N2; ++j) {
    const int b = m2[j];
    xm[j] = has_m ? x[b + m] : T(0);
    xp[j] = has_p ? x[b + m - 1] : T(0);
  }
  const uint32_t ft = first ? first[pos] : 0xffffffffu;
  if constexpr (R::kStatic) {
    // every row of this thread sits at element e: bottom rows of layer m, top rows of layer m-1.
    // The element is read as 16-byte chunks in entry order (one dwordx4 per lane, 1 KiB per
    // wavefront); row i accumulates over df2 in order in c[i].
    constexpr int V = chunk_entries<T>();
    constexpr int NP = N1 * N2, NC = (NP + V - 1) / V;
    typedef T VecT __attribute__((ext_vector_type(V)));
    const T *Ae = A + static_cast<size_t>(e >> 6) * gs + static_cast<size_t>(e & 63) * V;
    T c[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) c[i] = T(0);
    if constexpr (NP <= 64) {
      // small elements: every chunk in flight at once
#pragma unroll
      for (int ch = 0; ch < NC; ++ch) {
        bool need = false;
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int p = ch * V + q;
          if (p < NP) need = need || (R::top(p / N2) ? has_p : has_m);
        }
        VecT v = {};
        if (need) v = *reinterpret_cast<const VecT *>(Ae + static_cast<size_t>(ch) * 64 * V);
#pragma unroll
        for (int q = 0; q < V; ++q) {
          const int p = ch * V + q;
          if (p < NP) c[p / N2] = fma(v[q], R::top(p / N2) ? xp[p % N2] : xm[p % N2], c[p / N2]);
        }
      }
    } else {
      // large elements (order 1): one row at a time, loading the chunks that cover it; a chunk
      // shared by two rows is read twice, the second time from cache
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        if (!(R::top(i) ? has_p : has_m)) continue;
        const int c0 = (i * N2) / V, c1 = (i * N2 + N2 - 1) / V;
        T s = T(0);
#pragma unroll
        for (int ch = c0; ch <= c1; ++ch) {
          const VecT v = *reinterpret_cast<const VecT *>(Ae + static_cast<size_t>(ch) * 64 * V);
#pragma unroll
          for (int q = 0; q < V; ++q) {
            const int p = ch * V + q;
            if (p >= i * N2 && p < i * N2 + N2) s = fma(v[q], R::top(i) ? xp[p - i * N2] : xm[p - i * N2], s);
          }
        }
        c[i] = s;
      }
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      if (R::top(i)) continue;
      const int pt = R::partner(i);
      if (!has_m && pt < 0) continue;
      const T out = pt >= 0 ? c[i] + c[pt < 0 ? 0 : pt] : c[i];
      const size_t d = static_cast<size_t>(m1[i]) + m;
      if ((ft >> i) & 1u)
        y[d] = out;
      else
        y[d] += out;
    }
  } else {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      if (roles.top(i)) continue;
      const int pt = roles.partner(i);
      if (!has_m && pt < 0) continue;
      T acc = T(0);
      if (has_m) {
#pragma unroll
        for (int j = 0; j < N2; ++j) acc = fma(A[aidx<T>(gs, e, i * N2 + j)], xm[j], acc);
      }
      if (pt >= 0 && has_p) {
#pragma unroll
        for (int j = 0; j < N2; ++j) acc = fma(A[aidx<T>(gs, e, pt * N2 + j)], xp[j], acc);
      }
      const size_t d = static_cast<size_t>(m1[i]) + m;
      if ((ft >> i) & 1u)
        y[d] = acc;
      else
        y[d] += acc;
    }
  }
}

// Row r of element e times x at levels b[j] + lev, accumulated onto acc in df2 order.
template <typename T, int N2>
__device__ inline T row_dot(const T *__restrict__ A, size_t gs, long long e, int r, const T *__restrict__ x,
                            const int (&b)[N2], int lev, T acc) {
  for_row<T, N2>(A, gs, e, r, [&](int j, T a) { acc = fma(a, x[b[j] + lev], acc); });
  return acc;
}

// LMA apply in gather form, one launch for all colours.  Thread (column, level m) of the
// output space sums, over the cells of its column in colour order, the top-partner row of
// layer m-1 and then the bottom row of layer m: the order in which the reference's coloured
// loop adds into lhs(map1(df)+k) (matrix_vector_kernel_mod.F90:130-138, k outer).  Every
// output dof is written once, with no read-modify-write and no zero fill.  With DOT the
// kernel also leaves one partial sum of x.y per block (square operators: the CG's p.(A p)),
// which a one-block kernel adds in block order.  `skip` (a solver's done flag) stops it.
template <typename T, int N2, class R, bool DOT>
__global__ __launch_bounds__(kBlock) void k_lma_gather(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                       const T *__restrict__ x, T *__restrict__ y,
                                                       const int *__restrict__ map2, const int2 *__restrict__ cols,
                                                       const int *__restrict__ off, const int *__restrict__ ent,
                                                       int ncols, int nl, R roles, double *partials,
                                                       const int *skip, int xwin) {
  if (skip && *skip) return;
  const int H = nl + 1;
  const int lb = xcd_block(blockIdx.x, gridDim.x, xwin);
  const long long t = static_cast<long long>(lb) * kBlock + threadIdx.x;
  bool valid = t < static_cast<long long>(ncols) * H;
  const int col = valid ? static_cast<int>(t / H) : 0;
  const int m = valid ? static_cast<int>(t - static_cast<long long>(col) * H) : 0;
  int2 ci = make_int2(0, 0);
  if (valid) {
    ci = cols[col];
    if (m == nl && !ci.y) valid = false;
  }
  T acc = T(0);
  if (valid) {
    const bool has_m = m < nl, has_p = m > 0;
    const int q0 = off[col], q1 = off[col + 1];
    constexpr int KMAX = 4;  // cells around a vertex column of a quadrilateral mesh
    if (sizeof(T) == 8 && N2 <= 8 && q1 - q0 <= KMAX) {
      // FP64: the list's loads are issued together (three dependent rounds for the whole
      // column, not three per cell), then the rows are added in list order.  (For FP32 the
      // extra registers cost more occupancy than the shorter chain gains.)
      const int cnt = q1 - q0;
      int en[KMAX], eb[KMAX], b[KMAX][N2];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) en[k] = k < cnt ? ent[q0 + k] : 0;
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        const int pos = en[k] >> 5;
        eb[k] = k < cnt ? ebase[pos] : 0;
#pragma unroll
        for (int j = 0; j < N2; ++j) b[k][j] = k < cnt ? map2[static_cast<size_t>(pos) * N2 + j] : 0;
      }
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        if (k >= cnt) break;
        const int i = en[k] & 31;
        const int pt = roles.partner(i);
        const long long e = static_cast<long long>(eb[k]) + m;
        if (pt >= 0 && has_p) acc = row_dot<T, N2>(A, gs, e, pt, x, b[k], m - 1, acc);
        if (has_m) acc = row_dot<T, N2>(A, gs, e, i, x, b[k], m, acc);
      }
    } else {
      for (int q = q0; q < q1; ++q) {
        const int en = ent[q];
        const int pos = en >> 5, i = en & 31;
        const int pt = roles.partner(i);
        const long long e = static_cast<long long>(ebase[pos]) + m;
        int b[N2];
#pragma unroll
        for (int j = 0; j < N2; ++j) b[j] = map2[static_cast<size_t>(pos) * N2 + j];
        if (pt >= 0 && has_p) acc = row_dot<T, N2>(A, gs, e, pt, x, b, m - 1, acc);
        if (has_m) acc = row_dot<T, N2>(A, gs, e, i, x, b, m, acc);
      }
    }
    y[static_cast<size_t>(ci.x) + m] = acc;
  }
  if constexpr (DOT) {
    double v[1] = {valid ? to_d(acc) * to_d(x[static_cast<size_t>(ci.x) + m]) : 0.0};
    block_partials<1>(v, partials, lb);
  }
}

// ---- row-gathered operator copy --------------------------------------------------------------
// For the gather apply of a square operator whose dof columns each take rows from at most
// KMAX cells, the rows that thread t = (column, level m) reads are stored together: slot
// s = 2k + h holds, for the k-th cell of the column's list, its top-partner row of layer m-1
// (h = 0) or its bottom row of layer m (h = 1), each padded to whole chunks (RS entries).
// Address of entry j of slot s: (t/64)*gsr + ((s*RS + j)/V*64 + t%64)*V + (s*RS + j)%V, so a
// wavefront reads every chunk of its 64 threads as one contiguous block and the whole copy is
// streamed once in order, as the cell layout is by the colour-ordered kernel.  Rows that do
// not exist (m = 0 top, m = nlayers bottom, k beyond the list) hold zeros.
// (kRowsKmax and rows_rs: colour_apply.h)

template <typename T, int N2, class R>
__global__ __launch_bounds__(kBlock) void k_build_rows(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                       T *__restrict__ rows, size_t gsr, const int2 *__restrict__ cols,
                                                       const int *__restrict__ off, const int *__restrict__ ent,
                                                       int ncols, int nl, R roles) {
  constexpr int V = chunk_entries<T>(), RS = rows_rs<T, N2>();
  const int H = nl + 1;
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= static_cast<long long>(ncols) * H) return;
  const int col = static_cast<int>(t / H), m = static_cast<int>(t - static_cast<long long>(col) * H);
  const int q0 = off[col], cnt = off[col + 1] - q0;
  T *Rt = rows + static_cast<size_t>(t >> 6) * gsr + static_cast<size_t>(t & 63) * V;
  for (int k = 0; k < kRowsKmax; ++k) {
    int pos = 0, i = 0, pt = -1;
    if (k < cnt) {
      pos = ent[q0 + k] >> 5;
      i = ent[q0 + k] & 31;
      pt = roles.partner(i);
    }
    const long long e = static_cast<long long>(ebase[pos]) + m;
    for (int h = 0; h < 2; ++h) {
      const int r = h == 0 ? pt : i;
      const bool live = k < cnt && (h == 0 ? (pt >= 0 && m > 0) : m < nl);
      for (int j = 0; j < RS; ++j) {
        const int p = (2 * k + h) * RS + j;
        const T v = (live && j < N2) ? A[aidx<T>(gs, e, r * N2 + j)] : T(0);
        Rt[static_cast<size_t>(p / V) * 64 * V + (p % V)] = v;
      }
    }
  }
}

// Gather apply over the row-gathered copy of a lowest-order W0 operator: the same sums in the
// same order as k_lma_gather (per cell of the list: top-partner row of layer m-1, then bottom
// row of layer m).  A W0 cell's top vertex k+4 lies one level above its bottom vertex k in the
// same dof column (the copy is only built when the dofmap says so), so the x entries of a
// cell's two rows are its 4 columns at levels m-1, m and m+1, and each is loaded once.  The
// copy is read once per apply: its loads are nontemporal, so the stream does not evict x.
template <typename T, bool DOT>
__global__ __launch_bounds__(kBlock) void k_lma_gather_rows_w0(const T *__restrict__ rows, size_t gsr,
                                                               const T *__restrict__ x, T *__restrict__ y,
                                                               const int *__restrict__ map2, const int2 *__restrict__ cols,
                                                               const int *__restrict__ off, const int *__restrict__ ent,
                                                               int ncols, int nl, double *partials, const int *skip,
                                                               int xwin) {
  if (skip && *skip) return;
  constexpr int V = chunk_entries<T>(), RS = rows_rs<T, 8>();
  typedef T VecT __attribute__((ext_vector_type(V)));
  const int H = nl + 1;
  const int lb = xcd_block(blockIdx.x, gridDim.x, xwin);
  const long long t = static_cast<long long>(lb) * kBlock + threadIdx.x;
  bool valid = t < static_cast<long long>(ncols) * H;
  const int col = valid ? static_cast<int>(t / H) : 0;
  const int m = valid ? static_cast<int>(t - static_cast<long long>(col) * H) : 0;
  int2 ci = make_int2(0, 0);
  if (valid) {
    ci = cols[col];
    if (m == nl && !ci.y) valid = false;
  }
  T acc = T(0);
  if (valid) {
    const bool has_m = m < nl, has_p = m > 0;
    const int q0 = off[col], cnt = off[col + 1] - q0;
    const T *Rt = rows + static_cast<size_t>(t >> 6) * gsr + static_cast<size_t>(t & 63) * V;
#pragma unroll
    for (int k = 0; k < kRowsKmax; ++k) {
      if (k >= cnt) break;
      const int pos = ent[q0 + k] >> 5;
      int b[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = map2[static_cast<size_t>(pos) * 8 + j];
      T xl[3][4];  // x of the cell's 4 columns at levels m-1, m, m+1
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        xl[0][j] = has_p ? x[b[j] + m - 1] : T(0);
        xl[1][j] = x[b[j] + m];
        xl[2][j] = has_m ? x[b[j] + m + 1] : T(0);
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h == 0 ? !has_p : !has_m) continue;
#pragma unroll
        for (int c = 0; c < RS / V; ++c) {
          const VecT v = __builtin_nontemporal_load(
              reinterpret_cast<const VecT *>(Rt + static_cast<size_t>((2 * k + h) * (RS / V) + c) * 64 * V));
#pragma unroll
          for (int q = 0; q < V; ++q) {
            const int j = c * V + q;
            acc = fma(v[q], j < 4 ? xl[h][j] : xl[h + 1][j - 4], acc);
          }
        }
      }
    }
    y[static_cast<size_t>(ci.x) + m] = acc;
  }
  if constexpr (DOT) {
    double v[1] = {valid ? to_d(acc) * to_d(x[static_cast<size_t>(ci.x) + m]) : 0.0};
    block_partials<1>(v, partials, lb);
  }
}

// ---- symmetric-packed apply of lowest-order W0 operators ------------------------------------
// For operators whose every element matrix is symmetric (bit for bit), each element's upper
// triangle T(i, j), i <= j (36 entries, row-major) is enough.  Thread t = (cell position,
// segment of S levels) owns units m = seg*S .. seg*S+S-1 (m <= nlayers) of its cell: unit m needs
// the bottom rows of the layer-m element and the top rows of the layer-(m-1) element, so the
// thread walks up its segment keeping the previous triangle in registers.  Its copy holds S+1
// triangles: slot s is layer seg*S - 1 + s (slot 0 repeats the previous segment's last one).
// Entry k of slot s of thread t sits at (t/64)*gss + ((s*NCH + k/V)*64 + t%64)*V + k%V, NCH =
// 36/V chunks, so each wavefront load is one contiguous 1 KiB block.
__host__ __device__ constexpr int tri_index(int i, int j) { return i * 8 - i * (i - 1) / 2 + (j - i); }

template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_sym(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                      T *__restrict__ sym, size_t gss, int ncells, int nl, int S,
                                                      int nseg, uint32_t row_top, int *__restrict__ asym) {
  constexpr int V = chunk_entries<T>(), NCH = 36 / V;
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= static_cast<long long>(ncells) * nseg) return;
  const int pos = static_cast<int>(t / nseg), seg = static_cast<int>(t - static_cast<long long>(pos) * nseg);
  T *St = sym + static_cast<size_t>(t >> 6) * gss + static_cast<size_t>(t & 63) * V;
  bool mismatch = false;
  for (int s = 0; s <= S; ++s) {
    const int L = seg * S - 1 + s;
    const bool live = L >= 0 && L < nl;
    const long long e = static_cast<long long>(ebase[pos]) + (live ? L : 0);
    for (int i = 0; i < 8; ++i)
      for (int j = i; j < 8; ++j) {
        T v = T(0);
        if (live) {
          v = A[aidx<T>(gs, e + ((row_top >> i) & 1u), i * 8 + j)];
          const T w = A[aidx<T>(gs, e + ((row_top >> j) & 1u), j * 8 + i)];
          if constexpr (sizeof(T) == 8)
            mismatch = mismatch || (__double_as_longlong(v) != __double_as_longlong(w));
          else
            mismatch = mismatch || (__float_as_uint(v) != __float_as_uint(w));
        }
        const int k = tri_index(i, j);
        St[(static_cast<size_t>(s) * NCH + k / V) * 64 * V + k % V] = v;
      }
  }
  if (mismatch) *asym = 1;  // any thread; every writer stores the same value
}

// phase 1: the 4 outputs of every unit, z[(pos*4 + i)*(nl+1) + m] = c_i + cp_i with c_i the fma
// chain of bottom row i (layer m) and cp_i that of top row i+4 (layer m-1), exactly as
// k_lma_colour forms them
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sym_units(const T *__restrict__ sym, size_t gss, const T *__restrict__ x,
                                                      const int *__restrict__ map, T *__restrict__ z, int ncells,
                                                      int nl, int S, int nseg, const int *skip) {
  if (skip && *skip) return;
  constexpr int V = chunk_entries<T>(), NCH = 36 / V;
  typedef T VecT __attribute__((ext_vector_type(V)));
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= static_cast<long long>(ncells) * nseg) return;
  const int pos = static_cast<int>(t / nseg), seg = static_cast<int>(t - static_cast<long long>(pos) * nseg);
  const int H = nl + 1, m0 = seg * S;
  const T *St = sym + static_cast<size_t>(t >> 6) * gss + static_cast<size_t>(t & 63) * V;
  auto load_tri = [&](int s, T (&tri)[36]) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      const VecT v = __builtin_nontemporal_load(reinterpret_cast<const VecT *>(St + (static_cast<size_t>(s) * NCH + c) * 64 * V));
#pragma unroll
      for (int q = 0; q < V; ++q) tri[c * V + q] = v[q];
    }
  };
  int b[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = map[static_cast<size_t>(pos) * 8 + j];
  auto xat = [&](int j, int lev) { return (lev >= 0 && lev <= nl) ? x[b[j] + lev] : T(0); };
  // triangles of layers m-1 and m, and the next one loaded a step ahead; x at levels m-1, m,
  // m+1 and the next level, also a step ahead (the loads of step u+1 are in flight while step u
  // computes)
  T prev[36], cur[36], nxt[36], xw[3][4], xn[4];
  load_tri(0, prev);
  if (m0 < nl) load_tri(1, nxt);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    xw[0][j] = xat(j, m0 - 1);
    xw[1][j] = xat(j, m0);
    xw[2][j] = xat(j, m0 + 1);
    xn[j] = xat(j, m0 + 2);
  }
  T *zp = z + static_cast<size_t>(pos) * 4 * H;
  for (int u = 0; u < S; ++u) {
    const int m = m0 + u;
    if (m > nl) break;
    const bool has_m = m < nl, has_p = m > 0;
#pragma unroll
    for (int k = 0; k < 36; ++k) cur[k] = nxt[k];
    if (u + 1 < S && m + 1 < nl) load_tri(u + 2, nxt);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T c = T(0), cp = T(0);
      if (has_m) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          c = fma(cur[j >= i ? tri_index(i, j) : tri_index(j, i)], j < 4 ? xw[1][j] : xw[2][j - 4], c);
      }
      if (has_p) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          cp = fma(prev[j >= i + 4 ? tri_index(i + 4, j) : tri_index(j, i + 4)], j < 4 ? xw[0][j] : xw[1][j - 4], cp);
      }
      zp[static_cast<size_t>(i) * H + m] = c + cp;
    }
#pragma unroll
    for (int k = 0; k < 36; ++k) prev[k] = cur[k];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      xw[0][j] = xw[1][j];
      xw[1][j] = xw[2][j];
      xw[2][j] = xn[j];
      xn[j] = xat(j, m + 3);
    }
  }
}

// phase 1 with a compile-time segment: every triangle of the segment is loaded before any
// arithmetic, so all of a thread's matrix loads are in flight together (same sums as above)
template <typename T, int S>
__global__ __launch_bounds__(kBlock) void k_sym_units_s(const T *__restrict__ sym, size_t gss, const T *__restrict__ x,
                                                        const int *__restrict__ map, T *__restrict__ z, int ncells,
                                                        int nl, int nseg, const int *skip) {
  if (skip && *skip) return;
  constexpr int V = chunk_entries<T>(), NCH = 36 / V;
  typedef T VecT __attribute__((ext_vector_type(V)));
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= static_cast<long long>(ncells) * nseg) return;
  const int pos = static_cast<int>(t / nseg), seg = static_cast<int>(t - static_cast<long long>(pos) * nseg);
  const int H = nl + 1, m0 = seg * S;
  const T *St = sym + static_cast<size_t>(t >> 6) * gss + static_cast<size_t>(t & 63) * V;
  VecT tri[S + 1][NCH];
#pragma unroll
  for (int s = 0; s <= S; ++s)
#pragma unroll
    for (int c = 0; c < NCH; ++c)
      tri[s][c] = __builtin_nontemporal_load(reinterpret_cast<const VecT *>(St + (static_cast<size_t>(s) * NCH + c) * 64 * V));
  int b[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = map[static_cast<size_t>(pos) * 8 + j];
  T xs[S + 2][4];  // x at levels m0-1 .. m0+S
#pragma unroll
  for (int l = 0; l < S + 2; ++l)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int lev = m0 - 1 + l;
      xs[l][j] = (lev >= 0 && lev <= nl) ? x[b[j] + lev] : T(0);
    }
  auto at = [&](int s, int k) { return tri[s][k / V][k % V]; };
  T *zp = z + static_cast<size_t>(pos) * 4 * H;
#pragma unroll
  for (int u = 0; u < S; ++u) {
    const int m = m0 + u;
    if (m > nl) break;
    const bool has_m = m < nl, has_p = m > 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      T c = T(0), cp = T(0);
      if (has_m) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          c = fma(at(u + 1, j >= i ? tri_index(i, j) : tri_index(j, i)), j < 4 ? xs[u + 1][j] : xs[u + 2][j - 4], c);
      }
      if (has_p) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          cp = fma(at(u, j >= i + 4 ? tri_index(i + 4, j) : tri_index(j, i + 4)), j < 4 ? xs[u][j] : xs[u + 1][j - 4], cp);
      }
      zp[static_cast<size_t>(i) * H + m] = c + cp;
    }
  }
}

// One unit per thread (the default copy).  Cells take slots in cell-id order, so that the
// threads of a window of blocks work on neighbouring cells and share their x entries in one
// L2: unit t = k*(nl+1) + m is (k-th cell, level m), the cell at colour-major position
// slotpos[k].  Slot t of the copy holds the triangle of that cell's layer-m element (m < nl; the
// slot of m = nl is empty), entry q at (t/64)*gss + ((q/V)*64 + t%64)*V + q%V.  A unit's thread
// forms the bottom rows' chains c_i of its element (x at levels m, m+1) and the top rows' chains
// of the same element, which belong to unit m+1: the next lane takes them with one shuffle.
// Every chain is the colour kernel's (same entries, same x, same order); the unit's 4 outputs
// go to z[(k*4 + i)*(nl+1) + m].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_sym1(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                       const int *__restrict__ slotpos, T *__restrict__ sym, size_t gss,
                                                       int ncells, int nl, uint32_t row_top, int *__restrict__ asym) {
  constexpr int V = chunk_entries<T>();
  const int H = nl + 1;
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= static_cast<long long>(ncells) * H) return;
  const int k = static_cast<int>(t / H), m = static_cast<int>(t - static_cast<long long>(k) * H);
  T *St = sym + static_cast<size_t>(t >> 6) * gss + static_cast<size_t>(t & 63) * V;
  const bool live = m < nl;
  const long long e = static_cast<long long>(ebase[slotpos[k]]) + (live ? m : 0);
  bool mismatch = false;
  for (int i = 0; i < 8; ++i)
    for (int j = i; j < 8; ++j) {
      T v = T(0);
      if (live) {
        v = A[aidx<T>(gs, e + ((row_top >> i) & 1u), i * 8 + j)];
        const T w = A[aidx<T>(gs, e + ((row_top >> j) & 1u), j * 8 + i)];
        if constexpr (sizeof(T) == 8)
          mismatch = mismatch || (__double_as_longlong(v) != __double_as_longlong(w));
        else
          mismatch = mismatch || (__float_as_uint(v) != __float_as_uint(w));
      }
      const int q = tri_index(i, j);
      St[static_cast<size_t>(q / V) * 64 * V + q % V] = v;
    }
  if (mismatch) *asym = 1;
}

// cellcol[k]: the 0-based base dofs of the k-th cell's 4 bottom vertex columns.  A wavefront
// takes 63 new units and repeats, in lane 0, the previous wavefront's last unit (whose outputs
// it does not write), so that every unit's lower neighbour is the previous lane; s1_threads
// gives the launch's thread count.
inline long long s1_threads(long long units) { return (units + 62) / 63 * 64; }

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sym_units1(const T *__restrict__ sym, size_t gss, const T *__restrict__ x,
                                                       const int4 *__restrict__ cellcol, T *__restrict__ z, int ncells,
                                                       int nl, const int *skip, int xwin) {
  if (skip && *skip) return;
  constexpr int V = chunk_entries<T>(), NCH = 36 / V;
  typedef T VecT __attribute__((ext_vector_type(V)));
  const int H = nl + 1;
  const long long g = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x, xwin)) * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const long long t = (g >> 6) * 63 + lane - 1;
  const bool valid = t >= 0 && t < static_cast<long long>(ncells) * H;
  const int k = valid ? static_cast<int>(t / H) : 0;
  const int m = valid ? static_cast<int>(t - static_cast<long long>(k) * H) : 0;
  const bool has_m = valid && m < nl, has_p = valid && m > 0;
  const long long tc = valid ? t : 0;
  const T *St = sym + static_cast<size_t>(tc >> 6) * gss + static_cast<size_t>(tc & 63) * V;
  VecT tri[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c)
    tri[c] = has_m ? __builtin_nontemporal_load(reinterpret_cast<const VecT *>(St + static_cast<size_t>(c) * 64 * V))
                   : VecT{};
  int b[4] = {0, 0, 0, 0};
  if (valid) {
    const int4 bb = cellcol[k];
    b[0] = bb.x;
    b[1] = bb.y;
    b[2] = bb.z;
    b[3] = bb.w;
  }
  T x0[4], x1[4];  // x at levels m and m+1
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    x0[j] = valid ? x[b[j] + m] : T(0);
    x1[j] = has_m ? x[b[j] + m + 1] : T(0);
  }
  auto at = [&](int q) { return tri[q / V][q % V]; };
  T c[4], cn[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T s = T(0), sn = T(0);
    if (has_m) {
#pragma unroll
      for (int j = 0; j < 8; ++j) s = fma(at(j >= i ? tri_index(i, j) : tri_index(j, i)), j < 4 ? x0[j] : x1[j - 4], s);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        sn = fma(at(j >= i + 4 ? tri_index(i + 4, j) : tri_index(j, i + 4)), j < 4 ? x0[j] : x1[j - 4], sn);
    }
    c[i] = s;
    cn[i] = sn;
  }
  // the top rows of the layer-(m-1) element: the previous lane's unit (m-1 of the same cell)
  T cp[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) cp[i] = __shfl_up(cn[i], 1, 64);
  if (!valid || lane == 0) return;
  T *zp = z + static_cast<size_t>(k) * 4 * H;
#pragma unroll
  for (int i = 0; i < 4; ++i) zp[static_cast<size_t>(i) * H + m] = c[i] + (has_p ? cp[i] : T(0));
}

// phase 2: y at (column, level) = the units' outputs of the column's cells in colour order, the
// first written, the rest added (the colour kernel's first-touch write and += order).  With
// `zu` (columns of at most 4 cells) the column's units come as one int4 (cell*4 + i, or -1 when
// the column has fewer cells) beside its base dof, so a thread waits for two rounds of loads;
// otherwise its list entries are loaded before its z values (three rounds).  Returns x.y for
// the point (0 when t is past the end).
template <typename T>
__device__ inline double sym_gather_point(long long t, const T *__restrict__ z, const T *__restrict__ x,
                                          T *__restrict__ y, const int2 *__restrict__ cols,
                                          const int *__restrict__ off, const int *__restrict__ ent,
                                          const int4 *__restrict__ zu, int ncols, int nl) {
  const int H = nl + 1;
  bool valid = t < static_cast<long long>(ncols) * H;
  const int col = valid ? static_cast<int>(t / H) : 0;
  const int m = valid ? static_cast<int>(t - static_cast<long long>(col) * H) : 0;
  int2 ci = make_int2(0, 0);
  if (valid) {
    ci = cols[col];
    if (m == nl && !ci.y) valid = false;
  }
  if (!valid) return 0.0;
  T acc = T(0);
  if (zu) {
    const int4 u = zu[col];
    const int us[4] = {u.x, u.y, u.z, u.w};
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = us[q] >= 0 ? z[static_cast<size_t>(us[q]) * H + m] : T(0);
    acc = v[0];
#pragma unroll
    for (int q = 1; q < 4; ++q)
      if (us[q] >= 0) acc = acc + v[q];
  } else {
    const int q0 = off[col], cnt = off[col + 1] - q0;
    constexpr int KU = 4;
    int en[KU];
#pragma unroll
    for (int q = 0; q < KU; ++q) en[q] = q < cnt ? ent[q0 + q] : 0;
    T v[KU];
#pragma unroll
    for (int q = 0; q < KU; ++q)
      v[q] = q < cnt ? z[(static_cast<size_t>(en[q] >> 5) * 4 + (en[q] & 31)) * H + m] : T(0);
    acc = v[0];
#pragma unroll
    for (int q = 1; q < KU; ++q)
      if (q < cnt) acc = acc + v[q];
    for (int q = KU; q < cnt; ++q) {
      const int e2 = ent[q0 + q];
      acc = acc + z[(static_cast<size_t>(e2 >> 5) * 4 + (e2 & 31)) * H + m];
    }
  }
  y[static_cast<size_t>(ci.x) + m] = acc;
  return to_d(acc) * to_d(x[static_cast<size_t>(ci.x) + m]);
}

// one point per thread
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sym_gather(const T *__restrict__ z, const T *__restrict__ x,
                                                        T *__restrict__ y, const int2 *__restrict__ cols,
                                                        const int *__restrict__ off, const int *__restrict__ ent,
                                                        const int4 *__restrict__ zu, int ncols, int nl, const int *skip,
                                                        int xwin) {
  if (skip && *skip) return;
  const long long t = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x, xwin)) * kBlock + threadIdx.x;
  (void)sym_gather_point<T>(t, z, x, y, cols, off, ent, zu, ncols, nl);
}

// with the x.y partial sums of the CG's p.(A p): a fixed grid of kRedBlocks blocks walks the
// points (grid stride), so the next kernel can add the kRedBlocks partials itself
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sym_gather_dot(const T *__restrict__ z, const T *__restrict__ x,
                                                            T *__restrict__ y, const int2 *__restrict__ cols,
                                                            const int *__restrict__ off, const int *__restrict__ ent,
                                                            const int4 *__restrict__ zu, int ncols, int nl,
                                                            double *partials, const int *skip, int xwin) {
  if (skip && *skip) return;
  const int lb = xcd_block(blockIdx.x, gridDim.x, xwin);
  const long long total = static_cast<long long>(ncols) * (nl + 1);
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  double d[1] = {0.0};
  for (long long t = static_cast<long long>(lb) * kBlock + threadIdx.x; t < total; t += stride)
    d[0] += sym_gather_point<T>(t, z, x, y, cols, off, ent, zu, ncols, nl);
  block_partials<1>(d, partials, lb);
}

// ---- symmetric-packed order-1 W0 apply (sym27_apply.h) ----------------------------------
// Build: thread u = k*nl + m fills unit u's slot with the upper triangle of the k-th cell's
// layer-m element (rows of top dofs are stored one level up in the level-aligned layout) and
// raises *asym when an entry differs from its transpose in any bit.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_sym27(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                        const int *__restrict__ slotpos, T *__restrict__ sym, size_t gss,
                                                        long long units, int nl, uint32_t row_top,
                                                        int *__restrict__ asym) {
  constexpr int V = chunk_entries<T>(), NQ = s27_chunks<T>() * V;
  const long long u = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (u >= units) return;
  const int k = static_cast<int>(u / nl), m = static_cast<int>(u - static_cast<long long>(k) * nl);
  T *St = sym + static_cast<size_t>(u >> 6) * gss + static_cast<size_t>(u & 63) * V;
  const long long e = static_cast<long long>(ebase[slotpos[k]]) + m;
  bool mismatch = false;
  for (int q = 0; q < NQ; ++q) {
    T v = T(0);
    if (q < kS27Tri) {
      const int i = s27_row(q), j = s27_col(q);
      v = A[aidx<T>(gs, e + ((row_top >> i) & 1u), i * kS27 + j)];
      const T w = A[aidx<T>(gs, e + ((row_top >> j) & 1u), j * kS27 + i)];
      if constexpr (sizeof(T) == 8)
        mismatch = mismatch || (__double_as_longlong(v) != __double_as_longlong(w));
      else
        mismatch = mismatch || (__float_as_uint(v) != __float_as_uint(w));
    }
    St[static_cast<size_t>(q / V) * 64 * V + q % V] = v;
  }
  if (mismatch) *asym = 1;
}

// phase 1: L = NR/NRL lanes per unit, s27_threads(units) threads; xwin: the XCD block order
template <typename T, int NR, int NRL>
__global__ __launch_bounds__(kBlock) void k_sym27_units(const T *__restrict__ sym, size_t gss, const T *__restrict__ x,
                                                        const int *__restrict__ map2, const int *__restrict__ slotpos,
                                                        T *__restrict__ z, long long units, int nl, const int *skip,
                                                        int xwin) {
  if (skip && *skip) return;
  sym27_unit<T, NR, NRL>(sym, gss, x, map2, slotpos, z, units, nl,
                         static_cast<long long>(xcd_block(blockIdx.x, gridDim.x, xwin)) * kBlock + threadIdx.x);
}

// phase 2, one point per thread; DOT (one RHS): the block's partial sum of x.y as well
template <typename T, int NR, bool DOT>
__global__ __launch_bounds__(kBlock) void k_sym27_gather(const T *__restrict__ z, const T *__restrict__ x,
                                                         T *__restrict__ y, const int2 *__restrict__ cols,
                                                         const int4 *__restrict__ zu, int ncols, int nl,
                                                         double *partials, const int *skip) {
  if (skip && *skip) return;
  typedef typename RVec<T, NR>::type VR;
  VR acc;
  size_t dof = 0;
  const bool valid = sym27_gather_point<T, NR>(static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x, z, cols,
                                               zu, ncols, nl, acc, dof);
  if (valid) *reinterpret_cast<VR *>(y + dof * NR) = acc;
  if constexpr (DOT) {
    static_assert(NR == 1, "the fused dot product is for one right-hand side");
    double d[1] = {valid ? to_d(acc[0]) * to_d(x[dof]) : 0.0};
    block_partials<1>(d, partials);
  }
}

// Several right-hand sides over the row-gathered copy: each RHS accumulates exactly as
// k_lma_gather_rows_w0 (and so as k_lma_gather_mrhs), x and y hold NR values per dof, RHS
// fastest.  The matrix streams once for all right-hand sides.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void k_lma_gather_rows_w0_mrhs(const T *__restrict__ rows, size_t gsr,
                                                                    const T *__restrict__ x, T *__restrict__ y,
                                                                    const int *__restrict__ map2,
                                                                    const int2 *__restrict__ cols,
                                                                    const int *__restrict__ off,
                                                                    const int *__restrict__ ent, int ncols, int nl,
                                                                    int xwin) {
  rows_point_w0_mrhs<T, NR>(rows, gsr, x, y, map2, cols, off, ent, ncols, nl,
                            static_cast<long long>(xcd_block(blockIdx.x, gridDim.x, xwin)) * kBlock + threadIdx.x);
}

// ---- several right-hand sides ------------------------------------------------------------
// Multi-RHS vectors hold NR values per dof, RHS fastest: (dof, r) at dof*NR + r.  Every RHS
// is accumulated exactly as the one-RHS kernels do (same fma order), so column r of a
// multi-RHS apply equals the one-RHS apply of column r bit for bit.

// gather form, NR right-hand sides per thread (see k_lma_gather)
template <typename T, int N2, int NR, class R>
__global__ __launch_bounds__(kBlock) void k_lma_gather_mrhs(const T *__restrict__ A, size_t gs,
                                                            const int *__restrict__ ebase, const T *__restrict__ x,
                                                            T *__restrict__ y, const int *__restrict__ map2,
                                                            const int2 *__restrict__ cols, const int *__restrict__ off,
                                                            const int *__restrict__ ent, int ncols, int nl, R roles,
                                                            int xwin) {
  typedef typename RVec<T, NR>::type VR;
  const int H = nl + 1;
  const long long t = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x, xwin)) * kBlock + threadIdx.x;
  if (t >= static_cast<long long>(ncols) * H) return;
  const int col = static_cast<int>(t / H);
  const int m = static_cast<int>(t - static_cast<long long>(col) * H);
  const int2 ci = cols[col];
  if (m == nl && !ci.y) return;
  const bool has_m = m < nl, has_p = m > 0;
  VR acc = {};
  const int q1 = off[col + 1];
  for (int q = off[col]; q < q1; ++q) {
    const int en = ent[q];
    const int pos = en >> 5, i = en & 31;
    const int pt = roles.partner(i);
    const long long e = static_cast<long long>(ebase[pos]) + m;
    int b[N2];
#pragma unroll
    for (int j = 0; j < N2; ++j) b[j] = map2[static_cast<size_t>(pos) * N2 + j];
    if (pt >= 0 && has_p) row_dot_mrhs<T, N2, NR>(A, gs, e, pt, x, b, m - 1, acc);
    if (has_m) row_dot_mrhs<T, N2, NR>(A, gs, e, i, x, b, m, acc);
  }
  *reinterpret_cast<VR *>(y + (static_cast<size_t>(ci.x) + m) * NR) = acc;
}

// colour-ordered scatter form, NR right-hand sides per thread, one (cell, level) unit per thread
// (see colour_apply.h)
template <typename T, int N1, int N2, int NR, class R>
__global__ __launch_bounds__(kBlock) void k_lma_colour_mrhs(const T *__restrict__ A, size_t gs,
                                                            const int *__restrict__ ebase, const T *__restrict__ x,
                                                            T *__restrict__ y, const int *__restrict__ map1,
                                                            const int *__restrict__ map2,
                                                            const uint32_t *__restrict__ first, int pos0, int npos,
                                                            int nl, R roles, int xwin) {
  const int H = nl + 1;
  const int t = xcd_block(blockIdx.x, gridDim.x, xwin) * kBlock + threadIdx.x;
  if (t >= npos * H) return;
  const int lp = t / H;
  const int m = t - lp * H;
  const int pos = pos0 + lp;
  colour_unit_mrhs<T, N1, N2, NR>(A, gs, static_cast<long long>(ebase[pos]) + m, map1 + static_cast<size_t>(pos) * N1,
                                  map2 + static_cast<size_t>(pos) * N2, first[pos], x, y, m, nl, roles);
}

// Colour-ordered multi-RHS apply of a lowest-order W0 FP32 operator on the matrix cores (see
// mfma_apply.h): 64 (cell, level) units per 256-thread block, 4 lanes per unit.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_lma_colour_mrhs_mfma(const float *__restrict__ A, size_t gs,
                                                                 const int *__restrict__ ebase,
                                                                 const float *__restrict__ x, float *__restrict__ y,
                                                                 const int *__restrict__ map1,
                                                                 const int *__restrict__ map2,
                                                                 const uint32_t *__restrict__ first, int pos0,
                                                                 int npos, int nl, int xwin) {
  const int H = nl + 1;
  const long long u = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x, xwin)) * (kBlock / 4) + (threadIdx.x >> 2);
  const bool valid = u < static_cast<long long>(npos) * H;
  const int lp = valid ? static_cast<int>(u / H) : 0;
  const int m = valid ? static_cast<int>(u - static_cast<long long>(lp) * H) : 0;
  const int pos = pos0 + lp;
  mfma_unit_w0<NR>(A, gs, static_cast<long long>(ebase[pos]) + m, map1 + static_cast<size_t>(pos) * 8,
                   map2 + static_cast<size_t>(pos) * 8, first[pos], x, y, m, nl, valid);
}

// y = A^T x: thread (cell, level m) of the `from` space; entry (to_df i, from_df j) of layer k
// sits at element ebase + k + top_to(i)
template <typename T, int NT, int NF>
__global__ __launch_bounds__(kBlock) void k_transpose_colour(const T *__restrict__ A, size_t gs,
                                                             const int *__restrict__ ebase, const T *__restrict__ x,
                                                             T *__restrict__ y, const int *__restrict__ map_to,
                                                             const int *__restrict__ map_from,
                                                             const uint32_t *__restrict__ first, int pos0, int npos,
                                                             int nl, Roles roles, uint32_t row_top) {
  const int H = nl + 1;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= npos * H) return;
  const int lp = t / H;
  const int m = t - lp * H;
  const int pos = pos0 + lp;
  const int *mt = map_to + static_cast<size_t>(pos) * NT;
  const int *mf = map_from + static_cast<size_t>(pos) * NF;
  const bool has_m = m < nl, has_p = m > 0;
  const long long e = static_cast<long long>(ebase[pos]) + m;
  T xm[NT], xp[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    xm[i] = has_m ? x[mt[i] + m] : T(0);
    xp[i] = has_p ? x[mt[i] + m - 1] : T(0);
  }
  const uint32_t ft = first[pos];
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    if ((roles.top_mask >> j) & 1u) continue;
    const int pt = roles.partner[j];
    if (!has_m && pt < 0) continue;
    T acc = T(0);
    if (has_m) {
#pragma unroll
      for (int i = 0; i < NT; ++i) acc = fma(A[aidx<T>(gs, e + ((row_top >> i) & 1u), i * NF + j)], xm[i], acc);
    }
    if (pt >= 0 && has_p) {
#pragma unroll
      for (int i = 0; i < NT; ++i) acc = fma(A[aidx<T>(gs, e - 1 + ((row_top >> i) & 1u), i * NF + pt)], xp[i], acc);
    }
    const size_t d = static_cast<size_t>(mf[j]) + m;
    if ((ft >> j) & 1u)
      y[d] = acc;
    else
      y[d] += acc;
  }
}

// ---- Helmholtz composite y = M3 x + tau D diag(M2)^-1 D^T x in two kernels ------------------
// (W3 <- W2 lowest order: D holds 1 x 6 per element, M3 1 x 1.)  The first is the transpose in
// gather form with the division by m2 folded in: W2 dof (column, level m) adds its cells'
// contributions in colour order, the first written and the rest added, each the colour-
// ordered transpose's chain (row of layer m, then the top partner's of layer m-1), then divides
// by m2: the transpose followed by k_div_inplace, bit for bit.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_helm_transpose_div(const T *__restrict__ D, size_t gs,
                                                              const int *__restrict__ ebase, const T *__restrict__ x3,
                                                              const int *__restrict__ map3, const T *__restrict__ m2,
                                                              T *__restrict__ t, const int2 *__restrict__ cols,
                                                              const int *__restrict__ off, const int *__restrict__ ent,
                                                              int ncols, int nl, Roles roles) {
  const int H = nl + 1;
  const long long p = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= static_cast<long long>(ncols) * H) return;
  const int col = static_cast<int>(p / H), m = static_cast<int>(p - static_cast<long long>(col) * H);
  const int2 ci = cols[col];
  if (m == nl && !ci.y) return;
  const bool has_m = m < nl, has_p = m > 0;
  T total = T(0);
  const int q0 = off[col], q1 = off[col + 1];
  for (int q = q0; q < q1; ++q) {
    const int en = ent[q], pos = en >> 5, j = en & 31;
    const int pt = roles.partner[j];
    const long long e = static_cast<long long>(ebase[pos]) + m;
    const int b3 = map3[pos];
    T acc = T(0);
    if (has_m) acc = fma(D[aidx<T>(gs, e, j)], x3[b3 + m], acc);
    if (pt >= 0 && has_p) acc = fma(D[aidx<T>(gs, e - 1, pt)], x3[b3 + m - 1], acc);
    total = q == q0 ? acc : total + acc;
  }
  const size_t d = static_cast<size_t>(ci.x) + m;
  t[d] = total / m2[d];
}

// The composite's copy of D in transpose-gather order.  A table built for the composite holds
// each W2 column's (at most two) cells as one int4, {ebase, W3 dof * 8 + face} per cell in
// the column's colour order (the second ebase -1 when the column has one cell); the W2 lowest-
// order roles are fixed (face 4 the bottom, its partner the top face 5).  Per W2 dof (column,
// level m) the copy holds the two entries of D that its row of D^T takes (a side face: the
// column's first and second cell at layer m; a bottom/top face column: the bottom face at
// layer m and the top face at layer m-1), zero where there is none, so the transpose reads one
// coalesced pair per dof instead of one value out of each cell's 16-byte chunk of faces, and
// one level of dependent index loads where the gather lists take three.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_dtc(const T *__restrict__ D, size_t gs, const int2 *__restrict__ cols,
                                                     const int4 *__restrict__ tab, int ncols, int nl,
                                                     T *__restrict__ out) {
  const int H = nl + 1;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ncols * H) return;
  const int col = p / H, m = p - col * H;
  const int2 ci = cols[col];
  if (m == nl && !ci.y) return;
  const int4 c = tab[col];
  T a = T(0), b = T(0);
  if ((c.y & 7) == 4) {
    if (m < nl) a = D[aidx<T>(gs, static_cast<long long>(c.x) + m, 4)];
    if (m > 0) b = D[aidx<T>(gs, static_cast<long long>(c.x) + m - 1, 5)];
  } else {
    a = D[aidx<T>(gs, static_cast<long long>(c.x) + m, c.y & 7)];
    if (c.z >= 0) b = D[aidx<T>(gs, static_cast<long long>(c.z) + m, c.w & 7)];
  }
  const size_t d = static_cast<size_t>(ci.x) + m;
  out[2 * d] = a;
  out[2 * d + 1] = b;
}

// k_helm_transpose_div over that copy and the table: the same terms in the same order, bit for
// bit
template <typename T>
__global__ __launch_bounds__(kBlock) void k_helm_transpose_div_dtc(const T *__restrict__ dtc,
                                                                  const T *__restrict__ x3, const T *__restrict__ m2,
                                                                  T *__restrict__ t, const int2 *__restrict__ cols,
                                                                  const int4 *__restrict__ tab, int ncols, int nl) {
  using T2 = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const int H = nl + 1;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ncols * H) return;
  const int col = p / H, m = p - col * H;
  const int2 ci = cols[col];
  if (m == nl && !ci.y) return;
  const int4 c = tab[col];
  const size_t d = static_cast<size_t>(ci.x) + m;
  const T2 v = reinterpret_cast<const T2 *>(dtc)[d];
  const int b3 = c.y >> 3;
  T total;
  if ((c.y & 7) == 4) {
    T acc = T(0);
    if (m < nl) acc = fma(v.x, x3[b3 + m], acc);
    if (m > 0) acc = fma(v.y, x3[b3 + m - 1], acc);
    total = acc;
  } else {
    total = fma(v.x, x3[b3 + m], T(0));
    if (c.z >= 0) total = total + fma(v.y, x3[(c.w >> 3) + m], T(0));
  }
  t[d] = total / m2[d];
}

// the second: per W3 dof (cell, layer m) the divergence row (k_lma_colour's chain over the 6
// faces), M3's 1 x 1 product and y = M3 x + tau (D t) as k_axpy forms it
template <typename T>
__global__ __launch_bounds__(kBlock) void k_helm_dg_m3(const T *__restrict__ D, size_t gsd, const T *__restrict__ M3,
                                                      size_t gs3, const int *__restrict__ ebase,
                                                      const T *__restrict__ t, const int *__restrict__ map2,
                                                      const T *__restrict__ x3, const int *__restrict__ map3,
                                                      T *__restrict__ y, int ncells, int nl, T tau) {
  const long long u = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (u >= static_cast<long long>(ncells) * nl) return;
  const int pos = static_cast<int>(u / nl), m = static_cast<int>(u - static_cast<long long>(pos) * nl);
  const long long e = static_cast<long long>(ebase[pos]) + m;
  const int *b2 = map2 + static_cast<size_t>(pos) * 6;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) acc = fma(D[aidx<T>(gsd, e, j)], t[b2[j] + m], acc);
  const size_t d = static_cast<size_t>(map3[pos]) + m;
  const T y3 = fma(M3[aidx<T>(gs3, e, 0)], x3[d], T(0));
  y[d] = y3 + tau * acc;
}

// The composite's compact copy of D's rows and M3 in k_helm_dg_m3's thread order u = pos * nl
// + m: row j < 6 of D at out[j * nu + u], M3 at out[6 * nu + u].  The element layout pads a
// 1 x 6 element to two 16-byte chunks and a 1 x 1 element to one (each thread of the apply then
// pulls whole chunks for one value); the copy holds the values alone, coalesced per entry.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_build_drc(const T *__restrict__ D, size_t gsd, const T *__restrict__ M3,
                                                     size_t gs3, const int *__restrict__ ebase, int ncells, int nl,
                                                     T *__restrict__ out) {
  const int nu = ncells * nl;
  const int u = blockIdx.x * kBlock + threadIdx.x;
  if (u >= nu) return;
  const int pos = u / nl, m = u - pos * nl;
  const long long e = static_cast<long long>(ebase[pos]) + m;
#pragma unroll
  for (int j = 0; j < 6; ++j) out[static_cast<size_t>(j) * nu + u] = D[aidx<T>(gsd, e, j)];
  out[static_cast<size_t>(6) * nu + u] = M3[aidx<T>(gs3, e, 0)];
}

// k_helm_dg_m3 over that copy: the same products in the same order, bit for bit
template <typename T>
__global__ __launch_bounds__(kBlock) void k_helm_dg_m3_drc(const T *__restrict__ drc, const T *__restrict__ t,
                                                          const int *__restrict__ map2, const T *__restrict__ x3,
                                                          const int *__restrict__ map3, T *__restrict__ y, int ncells,
                                                          int nl, T tau) {
  const int nu = ncells * nl;
  const int u = blockIdx.x * kBlock + threadIdx.x;
  if (u >= nu) return;
  const int pos = u / nl, m = u - pos * nl;
  const int *b2 = map2 + static_cast<size_t>(pos) * 6;
  T acc = T(0);
#pragma unroll
  for (int j = 0; j < 6; ++j) acc = fma(drc[static_cast<size_t>(j) * nu + u], t[b2[j] + m], acc);
  const size_t d = static_cast<size_t>(map3[pos]) + m;
  const T y3 = fma(drc[static_cast<size_t>(6) * nu + u], x3[d], T(0));
  y[d] = y3 + tau * acc;
}

// assembled diagonal: sum over cells of A_k(df,df) (sci_mm_diagonal_kernel_mod.F90:87-92)
template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_diag_colour(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                        T *__restrict__ d, const int *__restrict__ map,
                                                        const uint32_t *__restrict__ first, int pos0, int npos, int nl,
                                                        Roles roles) {
  const int H = nl + 1;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= npos * H) return;
  const int lp = t / H;
  const int m = t - lp * H;
  const int pos = pos0 + lp;
  const int *mp = map + static_cast<size_t>(pos) * N;
  const uint32_t ft = first[pos];
  const long long e = static_cast<long long>(ebase[pos]) + m;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if ((roles.top_mask >> i) & 1u) continue;
    const int pt = roles.partner[i];
    if (m == nl && pt < 0) continue;
    T acc = T(0);
    if (m < nl) acc += A[aidx<T>(gs, e, i * N + i)];
    if (pt >= 0 && m > 0) acc += A[aidx<T>(gs, e, pt * N + pt)];
    const size_t o = static_cast<size_t>(mp[i]) + m;
    if ((ft >> i) & 1u)
      d[o] = acc;
    else
      d[o] += acc;
  }
}

// ---------------------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------------------
#define GRID_STRIDE(i, n)                                                                         \
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); \
       i += static_cast<long long>(gridDim.x) * blockDim.x)

template <typename T>
__global__ void k_set(T *v, long long n, T s) {
  GRID_STRIDE(i, n) v[i] = s;
}
template <typename T>
__global__ void k_axpy(T *__restrict__ y, const T *__restrict__ x, long long n, T a) {
  GRID_STRIDE(i, n) y[i] = y[i] + a * x[i];
}
template <typename T>
__global__ void k_aypx(T *__restrict__ y, const T *__restrict__ x, long long n, T a) {
  GRID_STRIDE(i, n) y[i] = a * y[i] + x[i];
}
template <typename T>
__global__ void k_axpby(T *__restrict__ z, T a, const T *__restrict__ x, T b, const T *__restrict__ y, long long n) {
  GRID_STRIDE(i, n) z[i] = a * x[i] + b * y[i];
}
template <typename T>
__global__ void k_scale(T *v, long long n, T a) {
  GRID_STRIDE(i, n) v[i] = a * v[i];
}
template <typename T>
__global__ void k_divide(T *__restrict__ z, const T *__restrict__ x, const T *__restrict__ d, long long n) {
  GRID_STRIDE(i, n) z[i] = x[i] / d[i];
}
template <typename T>
__global__ void k_axpy_mixed(double *__restrict__ y, double a, const T *__restrict__ x, long long n) {
  GRID_STRIDE(i, n) y[i] = fma(a, to_d(x[i]), y[i]);
}
// Streaming loops of the reduction kernels: 4 independent loads per vector in flight per
// thread before any arithmetic.
#define UNROLLED_STRIDE(n)                                                                   \
  const long long s_ = static_cast<long long>(gridDim.x) * blockDim.x;                     \
  for (long long i0_ = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i0_ < (n); i0_ += 4 * s_)

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(const T *__restrict__ x, const T *__restrict__ y, long long n,
                                                double *partials, unsigned int *counter, double *out) {
  double v[1] = {0.0};
  UNROLLED_STRIDE(n) {
    T a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0_ + u * s_;
      a[u] = i < n ? x[i] : T(0);
      b[u] = i < n ? y[i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) v[0] += to_d(a[u]) * to_d(b[u]);
  }
  block_partials<1>(v, partials);
}

// ---------------------------------------------------------------------------------------
// CG (cg_solve of components/science/source/solver/sci_iterative_solver_mod.f90:495-592)
// with scalars resident in SolverState
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ void k_cg_residual(T *__restrict__ r, const T *__restrict__ b, long long n) {
  GRID_STRIDE(i, n) r[i] = -r[i] + b[i];  // r%scale(-1); r%axpy(1,b)
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init_norm(const T *__restrict__ r, long long n, SolverState *st,
                                                         double *partials) {
  double v[1] = {0.0};
  UNROLLED_STRIDE(n) {
    T a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0_ + u * s_;
      a[u] = i < n ? r[i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) v[0] += to_d(a[u]) * to_d(a[u]);
  }
  block_partials<1>(v, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init_prec(const T *__restrict__ r, T *__restrict__ z, T *__restrict__ p,
                                                         const T *__restrict__ diag, long long n, SolverState *st,
                                                         double *partials) {
  if (st->done) return;
  double v[1] = {0.0};
  GRID_STRIDE(i, n) {
    const T zi = diag ? r[i] / diag[i] : r[i];
    z[i] = zi;
    p[i] = zi;
    v[0] += to_d(r[i]) * to_d(zi);
  }
  block_partials<1>(v, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_dot_pz(const T *__restrict__ p, const T *__restrict__ z, long long n,
                                                      SolverState *st, double *partials) {
  if (st->done) return;
  double v[1] = {0.0};
  UNROLLED_STRIDE(n) {
    T a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0_ + u * s_;
      a[u] = i < n ? p[i] : T(0);
      b[u] = i < n ? z[i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) v[0] += to_d(a[u]) * to_d(b[u]);
  }
  block_partials<1>(v, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                                                      const T *__restrict__ p, const T *__restrict__ diag,
                                                      long long n, SolverState *st, double *partials) {
  if (st->done) return;
  const T a = static_cast<T>(st->alpha);
  const T ma = static_cast<T>(-st->alpha);
  double v[2] = {0.0, 0.0};
  UNROLLED_STRIDE(n) {
    T xv[4], pv[4], rv[4], zv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0_ + u * s_;
      const bool ok = i < n;
      xv[u] = ok ? x[i] : T(0);
      pv[u] = ok ? p[i] : T(0);
      rv[u] = ok ? r[i] : T(0);
      zv[u] = ok ? z[i] : T(0);
      dv[u] = (ok && diag) ? diag[i] : T(1);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0_ + u * s_;
      if (i >= n) continue;
      x[i] = xv[u] + a * pv[u];              // x%axpy(alpha,p)
      const T ri = rv[u] + ma * zv[u];       // r%axpy(-alpha,z)
      r[i] = ri;
      const T zi = diag ? ri / dv[u] : ri;   // prec%apply(r,z)
      z[i] = zi;
      v[0] += to_d(ri) * to_d(ri);
      v[1] += to_d(ri) * to_d(zi);
    }
  }
  block_partials<2>(v, partials);
}

// Finishing step of every reduction: sum the block partials in block order, then update the
// solver scalars (cg_solve's scalar arithmetic, sci_iterative_solver_mod.f90:521-575).
enum Epilogue { kEpiDot = 0, kEpiInitNorm, kEpiInitPrec, kEpiDotPz };

// The sum of a reduction's block partials, the same in every block that forms it: thread t
// adds partials t, t + kBlock, ... in kChains independent chains (chain u takes every
// kChains-th of them), so the loads overlap; the chains are added in order, then the block's
// threads in block_sum's fixed tree.  Every thread of the block gets the total.
template <int NV>
__device__ inline void reduce_partials(const double *__restrict__ partials, int nblocks, double (&tot)[NV]) {
  constexpr int kChains = 8;
  double c[kChains][NV];
#pragma unroll
  for (int u = 0; u < kChains; ++u)
#pragma unroll
    for (int q = 0; q < NV; ++q) c[u][q] = 0.0;
  int b = threadIdx.x;
  for (; b + (kChains - 1) * kBlock < nblocks; b += kChains * kBlock)
#pragma unroll
    for (int u = 0; u < kChains; ++u)
#pragma unroll
      for (int q = 0; q < NV; ++q) c[u][q] += partials[(b + u * kBlock) * NV + q];
#pragma unroll
  for (int u = 0; u < kChains - 1; ++u)
    if (b + u * kBlock < nblocks)
#pragma unroll
      for (int q = 0; q < NV; ++q) c[u][q] += partials[(b + u * kBlock) * NV + q];
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    s[q] = c[0][q];
#pragma unroll
    for (int u = 1; u < kChains; ++u) s[q] += c[u][q];
  }
  block_sum<NV>(s, tot);
}

template <int NV, int EPI>
__global__ __launch_bounds__(kBlock) void k_finish(const double *__restrict__ partials, int nblocks, SolverState *st,
                                                   double *out, int par) {
  if (EPI == kEpiInitPrec || EPI == kEpiDotPz) {
    if (st->done) return;
  }
  double tot[NV];
  reduce_partials<NV>(partials, nblocks, tot);
  if (threadIdx.x != 0) return;
  if (EPI == kEpiDot) {
    *out = tot[0];
  } else if (EPI == kEpiInitNorm) {
    const double nrm = sqrt(tot[0]);
    st->r_nrm0 = nrm;
    st->r_nrm = nrm;
    st->iter = 0;
    if (nrm < 2.220446049250313e-16) {  // ||r0|| < EPS: converged in 0 iterations (:527-532)
      st->done = 1;
      st->converged = 1;
    }
    if (nrm != nrm) {
      st->done = 1;
      st->status = LFRIC_NAN;
    }
  } else if (EPI == kEpiInitPrec) {
    st->rz = tot[0];
    st->rzb[0] = tot[0];
  } else if (EPI == kEpiDotPz) {
    const double pz = tot[0];
    st->pz = pz;
    if (pz != pz) {
      st->status = LFRIC_NAN;
      st->done = 1;
    } else if (pz == 0.0) {
      st->status = LFRIC_BREAKDOWN;
      st->done = 1;
    } else {
      st->alpha = st->rzb[par] / pz;  // alpha = <r,z> / <p,A.p>
    }
  }
}

// CG's update with the p.(A p) finish folded in: every block adds the npz partials of the apply
// itself (reduce_partials: the same total in every block) and forms alpha = rz / pz; block 0
// records it (and a NaN or zero pz, which stops the solve in every block alike).  par selects
// the iteration's rz slot (the host counts the iterations it enqueues).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update_fold(T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                                                           const T *__restrict__ p, const T *__restrict__ diag,
                                                           long long n, SolverState *st,
                                                           const double *__restrict__ pzp, int npz,
                                                           double *partials, int par) {
  if (st->done) return;
  double tp[1];
  reduce_partials<1>(pzp, npz, tp);
  const double pz = tp[0];
  if (pz != pz || pz == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->pz = pz;
      st->status = pz != pz ? LFRIC_NAN : LFRIC_BREAKDOWN;
      st->done = 1;
    }
    return;
  }
  const double alpha = st->rzb[par] / pz;  // alpha = <r,z> / <p,A.p>
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pz = pz;
    st->alpha = alpha;
  }
  const T a = static_cast<T>(alpha);
  const T ma = static_cast<T>(-alpha);
  double v[2] = {0.0, 0.0};
  UNROLLED_STRIDE(n) {
    T xv[4], pv[4], rv[4], zv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0_ + u * s_;
      const bool ok = i < n;
      xv[u] = ok ? x[i] : T(0);
      pv[u] = ok ? p[i] : T(0);
      rv[u] = ok ? r[i] : T(0);
      zv[u] = ok ? z[i] : T(0);
      dv[u] = (ok && diag) ? diag[i] : T(1);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long long i = i0_ + u * s_;
      if (i >= n) continue;
      x[i] = xv[u] + a * pv[u];              // x%axpy(alpha,p)
      const T ri = rv[u] + ma * zv[u];       // r%axpy(-alpha,z)
      r[i] = ri;
      const T zi = diag ? ri / dv[u] : ri;   // prec%apply(r,z)
      z[i] = zi;
      v[0] += to_d(ri) * to_d(ri);
      v[1] += to_d(ri) * to_d(zi);
    }
  }
  block_partials<2>(v, partials);
}

// CG's aypx with the update's finish folded in: every block adds the update's partials (r.r,
// r.z), tests convergence as cg_solve does (monitor on: ||r|| / ||r0|| <= r_tol or ||r|| <=
// a_tol) and, when the solve goes on, forms beta = rz_new / rz and p = beta p + z.  Block 0
// keeps the scalars: the iteration count, ||r||, the next rz slot, the done flag.  A block that
// starts after block 0 has set done skips p, which the finished solve no longer reads.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_aypx_fold(T *__restrict__ p, const T *__restrict__ z, long long n,
                                                         SolverState *st, const double *__restrict__ partials,
                                                         int nblocks, int par) {
  if (st->done) return;
  double tot[2];
  reduce_partials<2>(partials, nblocks, tot);
  const double rn = sqrt(tot[0]);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (rn != rn) {
    if (lead) {
      st->iter += 1;
      st->r_nrm = rn;
      st->status = LFRIC_NAN;
      st->done = 1;
    }
    return;
  }
  if (st->monitor && (rn / st->r_nrm0 <= st->r_tol || rn <= st->a_tol)) {
    if (lead) {
      st->iter += 1;
      st->r_nrm = rn;
      st->red0 = tot[0];
      st->red1 = tot[1];
      st->converged = 1;
      st->done = 1;
    }
    return;
  }
  const double rz_new = tot[1];
  const double beta = rz_new / st->rzb[par];
  const T b = static_cast<T>(beta);
  GRID_STRIDE(i, n) p[i] = b * p[i] + z[i];  // p%aypx(beta,z)
  if (lead) {
    st->iter += 1;
    st->r_nrm = rn;
    st->red0 = tot[0];
    st->red1 = tot[1];
    st->beta = beta;
    st->rz = rz_new;
    st->rzb[par ^ 1] = rz_new;
    if (st->iter >= st->max_iter) st->done = 1;
  }
}

template <typename T>
__global__ void k_cg_aypx(T *__restrict__ p, const T *__restrict__ z, long long n, SolverState *st) {
  if (st->done) return;
  const T b = static_cast<T>(st->beta);
  GRID_STRIDE(i, n) p[i] = b * p[i] + z[i];  // p%aypx(beta,z)
}

// ---------------------------------------------------------------------------------------
}  // namespace

Roles roles_of(const FunctionSpace &fs) {
  Roles r{};
  r.top_mask = fs.top_mask;
  for (int i = 0; i < 32; ++i) r.partner[i] = static_cast<int8_t>(i < fs.ndf ? fs.partner[i] : -1);
  return r;
}

// ----- launchers ---------------------------------------------------------------------------
// ---- dg_inc apply and boundary conditions ------------------------------------------------------
// dg_inc_matrix_vector_code (components/science/source/kernel/algebra/
// dg_inc_matrix_vector_kernel_mod.F90:69-142): y(map1(df)+k) += sum over df2, in df2 order, of
// A(ik+k,df,df2) x(map2(df2)+k).  The output space is discontinuous (each output dof belongs to
// one cell), so thread (cell position, layer k) owns its outputs: no colours, no atomics.  Each
// output's chain starts from the value y holds, as the reference's statement does.
template <typename T, int N1, int N2>
__global__ __launch_bounds__(kBlock) void k_dg_inc(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                   const T *__restrict__ x, T *__restrict__ y,
                                                   const int *__restrict__ map1, const int *__restrict__ map2,
                                                   int npos, int nl, int n1rt, int n2rt) {
  const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= static_cast<long long>(npos) * nl) return;
  const int pos = static_cast<int>(t / nl), k = static_cast<int>(t - static_cast<long long>(pos) * nl);
  const int n1 = N1 > 0 ? N1 : n1rt, n2 = N2 > 0 ? N2 : n2rt;
  const long long e = static_cast<long long>(ebase[pos]) + k;
  for (int i = 0; i < n1; ++i) {
    const size_t d = static_cast<size_t>(map1[static_cast<size_t>(pos) * n1 + i]) + k;
    T acc = y[d];
    for (int j = 0; j < n2; ++j)
      acc = fma(A[aidx<T>(gs, e, i * n2 + j)], x[static_cast<size_t>(map2[static_cast<size_t>(pos) * n2 + j]) + k], acc);
    y[d] = acc;
  }
}

// enforce_bc_code (components/science/source/kernel/algebra/sci_enforce_bc_kernel_mod.F90:63-91):
// field(map(df)) *= boundary_value(df,1) and field(map(df)+nlayers-1) *= boundary_value(df,2) in
// every cell column.  The flags are 0 or 1; a product with 1 leaves the value as it is, so only
// the dofs flagged 0 are touched (one thread per cell position, the bits of `bottom` / `top`).
// Two cells that flag one shared dof both write x*0 of the same x: the result does not depend
// on their order, because (x*0)*0 == x*0.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_enforce_bc(T *__restrict__ y, const int *__restrict__ map, int npos, int ndf,
                                                       int nl, uint32_t bottom, uint32_t top, const int *skip) {
  if (skip && *skip) return;
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= npos) return;
  for (int df = 0; df < ndf; ++df) {
    const size_t b = static_cast<size_t>(map[static_cast<size_t>(pos) * ndf + df]);
    if ((bottom >> df) & 1u) y[b] = y[b] * T(0);
    if ((top >> df) & 1u) y[b + nl - 1] = y[b + nl - 1] * T(0);
  }
}

#define LAUNCH(kernel, grid, ...) hipLaunchKernelGGL((kernel), dim3(grid), dim3(kBlock), 0, cur_stream(), __VA_ARGS__)
#define CHECK_LAUNCH() LF_HIP(hipGetLastError())

int launch_gen_spd(lfric_op *op, unsigned long long seed, int scaled, double shift) {
  const FunctionSpace &f = *op->to;
  LF_CHECK(op->to == op->from || (f.ndf == op->from->ndf), LFRIC_E_ARG, "SPD generator needs a square operator");
  const long long nelem = static_cast<long long>(f.ncells) * f.nlayers;
  int *d_cop = nullptr;
  size_t bytes = sizeof(int) * f.cell_of_pos.size();
  if (int rc = dmalloc(reinterpret_cast<void **>(&d_cop), bytes)) return rc;
  LF_HIP(hipMemcpyAsync(d_cop, f.cell_of_pos.data(), bytes, hipMemcpyHostToDevice, cur_stream()));
  const int grid = grid_for(nelem, kBlock / 64);
  int rc = LFRIC_OK;
  auto go = [&](auto tag, auto ndfc) {
    using T = decltype(tag);
    constexpr int N = decltype(ndfc)::value;
    LAUNCH((k_gen_spd<T, N>), grid, static_cast<T *>(op->d), op->gs, f.d_ebase, d_cop, f.nlayers, nelem, seed,
           scaled, shift, op->row_top);
  };
  switch (f.ndf) {
    case 1: op->precision == LFRIC_R64 ? go(double{}, std::integral_constant<int, 1>{}) : go(float{}, std::integral_constant<int, 1>{}); break;
    case 6: op->precision == LFRIC_R64 ? go(double{}, std::integral_constant<int, 6>{}) : go(float{}, std::integral_constant<int, 6>{}); break;
    case 8: op->precision == LFRIC_R64 ? go(double{}, std::integral_constant<int, 8>{}) : go(float{}, std::integral_constant<int, 8>{}); break;
    case 27: op->precision == LFRIC_R64 ? go(double{}, std::integral_constant<int, 27>{}) : go(float{}, std::integral_constant<int, 27>{}); break;
    default: rc = fail(LFRIC_E_UNSUPPORTED, "generator: unsupported ndf");
  }
  hipError_t e = hipGetLastError();
  LF_HIP(hipStreamSynchronize(cur_stream()));
  dfree(d_cop, bytes);
  if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
  return rc;
}

int launch_gen_uniform(lfric_op *op, unsigned long long seed, unsigned long long stream, double lo, double hi,
                       const double *col_sign_host) {
  const FunctionSpace &f = *op->to;
  const int n1 = op->to->ndf, n2 = op->from->ndf;
  const long long nelem = static_cast<long long>(f.ncells) * f.nlayers;
  int *d_cop = nullptr;
  double *d_sign = nullptr;
  size_t bytes = sizeof(int) * f.cell_of_pos.size();
  if (int rc = dmalloc(reinterpret_cast<void **>(&d_cop), bytes)) return rc;
  if (col_sign_host) {
    if (int rc = dmalloc(reinterpret_cast<void **>(&d_sign), sizeof(double) * n2)) return rc;
    LF_HIP(hipMemcpyAsync(d_sign, col_sign_host, sizeof(double) * n2, hipMemcpyHostToDevice, cur_stream()));
  }
  LF_HIP(hipMemcpyAsync(d_cop, f.cell_of_pos.data(), bytes, hipMemcpyHostToDevice, cur_stream()));
  const int grid = grid_for(nelem);
  if (op->precision == LFRIC_R64)
    LAUNCH(k_gen_uniform<double>, grid, static_cast<double *>(op->d), op->gs, f.d_ebase, d_cop, f.nlayers, nelem, n1,
           n2, seed, stream, lo, hi, d_sign, op->row_top);
  else
    LAUNCH(k_gen_uniform<float>, grid, static_cast<float *>(op->d), op->gs, f.d_ebase, d_cop, f.nlayers, nelem, n1,
           n2, seed, stream, lo, hi, d_sign, op->row_top);
  hipError_t e = hipGetLastError();
  LF_HIP(hipStreamSynchronize(cur_stream()));
  dfree(d_cop, bytes);
  if (d_sign) dfree(d_sign, sizeof(double) * n2);
  if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
  return LFRIC_OK;
}

int launch_vec_uniform(lfric_vec *v, unsigned long long seed, unsigned long long stream) {
  const int grid = grid_for(v->n, kBlock, 65536);
  if (v->precision == LFRIC_R64)
    LAUNCH(k_vec_uniform<double>, grid, static_cast<double *>(v->d), v->n, seed, stream);
  else
    LAUNCH(k_vec_uniform<float>, grid, static_cast<float *>(v->d), v->n, seed, stream);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

static int with_cell_of_pos(const FunctionSpace &f, int **d_cop, size_t &bytes) {
  bytes = sizeof(int) * f.cell_of_pos.size();
  if (int rc = dmalloc(reinterpret_cast<void **>(d_cop), bytes)) return rc;
  LF_HIP(hipMemcpyAsync(*d_cop, f.cell_of_pos.data(), bytes, hipMemcpyHostToDevice, cur_stream()));
  return LFRIC_OK;
}

template <typename S>
static int op_from_reference_t(lfric_op *op, const S *d_ref) {
  const FunctionSpace &f = *op->to;
  int *d_cop = nullptr;
  size_t bytes = 0;
  if (int rc = with_cell_of_pos(f, &d_cop, bytes)) return rc;
  const long long nelem = static_cast<long long>(f.ncells) * f.nlayers;
  const int n1 = op->to->ndf, n2 = op->from->ndf;
  if (op->precision == LFRIC_R64)
    LAUNCH((k_from_reference<double, S>), grid_for(nelem), static_cast<double *>(op->d), op->gs, f.d_ebase, d_ref,
           d_cop, f.nlayers, nelem, n1, n2, op->row_top);
  else
    LAUNCH((k_from_reference<float, S>), grid_for(nelem), static_cast<float *>(op->d), op->gs, f.d_ebase, d_ref,
           d_cop, f.nlayers, nelem, n1, n2, op->row_top);
  hipError_t e = hipGetLastError();
  LF_HIP(hipStreamSynchronize(cur_stream()));
  dfree(d_cop, bytes);
  if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
  return LFRIC_OK;
}

template <typename D>
static int op_to_reference_t(const lfric_op *op, D *d_ref) {
  const FunctionSpace &f = *op->to;
  int *d_cop = nullptr;
  size_t bytes = 0;
  if (int rc = with_cell_of_pos(f, &d_cop, bytes)) return rc;
  const long long nelem = static_cast<long long>(f.ncells) * f.nlayers;
  const int n1 = op->to->ndf, n2 = op->from->ndf;
  if (op->precision == LFRIC_R64)
    LAUNCH((k_to_reference<double, D>), grid_for(nelem), static_cast<const double *>(op->d), op->gs, f.d_ebase,
           d_ref, d_cop, f.nlayers, nelem, n1, n2, op->row_top);
  else
    LAUNCH((k_to_reference<float, D>), grid_for(nelem), static_cast<const float *>(op->d), op->gs, f.d_ebase,
           d_ref, d_cop, f.nlayers, nelem, n1, n2, op->row_top);
  hipError_t e = hipGetLastError();
  LF_HIP(hipStreamSynchronize(cur_stream()));
  dfree(d_cop, bytes);
  if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
  return LFRIC_OK;
}

int launch_op_from_reference(lfric_op *op, const void *d_ref, int ref_precision) {
  return ref_precision == LFRIC_R64 ? op_from_reference_t(op, static_cast<const double *>(d_ref))
                                    : op_from_reference_t(op, static_cast<const float *>(d_ref));
}

int launch_op_to_reference(const lfric_op *op, void *d_ref, int ref_precision) {
  return ref_precision == LFRIC_R64 ? op_to_reference_t(op, static_cast<double *>(d_ref))
                                    : op_to_reference_t(op, static_cast<float *>(d_ref));
}

template <typename T>
static int invert_local_t(const lfric_op *A, lfric_op *Ainv) {
  const FunctionSpace &f = *A->to;
  const long long nelem = static_cast<long long>(f.ncells) * f.nlayers;
  const int n = f.ndf;
  const T *a = static_cast<const T *>(A->d);
  T *b = static_cast<T *>(Ainv->d);
  switch (n) {
    case 1: LAUNCH((k_invert_local<T, 1>), grid_for(nelem), a, b, A->gs, f.d_ebase, f.nlayers, nelem, A->row_top); break;
    case 2: LAUNCH((k_invert_local<T, 2>), grid_for(nelem), a, b, A->gs, f.d_ebase, f.nlayers, nelem, A->row_top); break;
    case 6: LAUNCH((k_invert_local<T, 6>), grid_for(nelem), a, b, A->gs, f.d_ebase, f.nlayers, nelem, A->row_top); break;
    case 8: LAUNCH((k_invert_local<T, 8>), grid_for(nelem), a, b, A->gs, f.d_ebase, f.nlayers, nelem, A->row_top); break;
    default: {
      // workspace for at most 2^20 elements at a time
      const long long chunk = std::min<long long>(nelem, 1LL << 20);
      const size_t bytes = sizeof(T) * static_cast<size_t>(n * n + 2 * n) * static_cast<size_t>(chunk);
      void *work = nullptr;
      if (int rc = dmalloc(&work, bytes)) return rc;
      for (long long e0 = 0; e0 < nelem; e0 += chunk) {
        const long long cnt = std::min(chunk, nelem - e0);
        LAUNCH(k_invert_local_ws<T>, grid_for(cnt), a, b, A->gs, f.d_ebase, f.nlayers, e0, cnt, n, A->row_top,
               static_cast<T *>(work));
      }
      const hipError_t e = hipGetLastError();
      const hipError_t es = hipStreamSynchronize(cur_stream());
      dfree(work, bytes);
      if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
      if (es != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(es));
      return LFRIC_OK;
    }
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

// Gauss-Legendre points and weights on [0, 1]: the Newton iteration of quadrature_rule in
// infrastructure/source/quadrature/quadrature_rule_gaussian_mod.F90:35-94 (each root's
// iteration starts from the previous root; capped here at 100 steps), mapped from [-1, 1].
static void gauss_rule(int n, double *pts, double *wts) {
  const double pi = 4.0 * std::atan(1.0), eps = 2.220446049250313e-16;
  double z1 = 0.0;
  for (int i = 1; i <= (n + 1) / 2; ++i) {
    double z = std::cos(pi * (i - 0.25) / (n + 0.5)), pp = 1.0;
    for (int it = 0; it < 100 && std::fabs(z - z1) > eps; ++it) {
      double p1 = 1.0, p2 = 0.0, p3;
      for (int j = 1; j <= n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    }
    pts[i - 1] = -z;
    pts[n - i] = z;
    wts[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
    wts[n - i] = wts[i - 1];
  }
  for (int i = 0; i < n; ++i) {
    pts[i] = 0.5 * (pts[i] + 1.0);
    wts[i] = 0.5 * wts[i];
  }
}

// Quadrature tables of the W0 trilinear basis (reference-cube vertices SWB, SEB, NEB, NWB, SWT,
// SET, NET, NWT, components of the reference element in reference_element_mod.F90:1244-1251):
// values [8][NP], derivatives [3][8][NP], horizontal weights [nh] (x_i with y_j, y fastest, as
// quadrature_xyoz_mod.F90:235-243), vertical weights [nq]; point q = kz*nh + i*nq + j.
static std::vector<double> w0_tables(int nq) {
  std::vector<double> p(nq), w(nq);
  gauss_rule(nq, p.data(), w.data());
  const int nh = nq * nq, np = nh * nq;
  static const int vx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, vy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, vz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  std::vector<double> t(static_cast<size_t>(32 * np + nh + nq));
  for (int kz = 0; kz < nq; ++kz)
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j) {
        const int q = kz * nh + i * nq + j;
        const double x = p[i], y = p[j], z = p[kz];
        for (int df = 0; df < 8; ++df) {
          const double lx = vx[df] ? x : 1.0 - x, ly = vy[df] ? y : 1.0 - y, lz = vz[df] ? z : 1.0 - z;
          const double dx = vx[df] ? 1.0 : -1.0, dy = vy[df] ? 1.0 : -1.0, dz = vz[df] ? 1.0 : -1.0;
          t[static_cast<size_t>(df) * np + q] = lx * ly * lz;
          t[static_cast<size_t>(8 * np) + static_cast<size_t>(0 * 8 + df) * np + q] = dx * ly * lz;
          t[static_cast<size_t>(8 * np) + static_cast<size_t>(1 * 8 + df) * np + q] = lx * dy * lz;
          t[static_cast<size_t>(8 * np) + static_cast<size_t>(2 * 8 + df) * np + q] = lx * ly * dz;
        }
      }
  for (int i = 0; i < nq; ++i)
    for (int j = 0; j < nq; ++j) t[static_cast<size_t>(32 * np) + i * nq + j] = w[i] * w[j];
  for (int kz = 0; kz < nq; ++kz) t[static_cast<size_t>(32 * np + nh) + kz] = w[kz];
  return t;
}

int launch_assemble_mass_w0(lfric_op *op, const double *chi1, const double *chi2, const double *chi3, int nqp) {
  const FunctionSpace &f = *op->to;
  bool w0 = op->to == op->from && f.ndf == 8 && f.top_mask == 0xF0u;
  for (int i = 0; w0 && i < 4; ++i) w0 = f.partner[i] == i + 4;
  LF_CHECK(w0, LFRIC_E_UNSUPPORTED, "assemble_mass_w0: a square lowest-order W0 operator is needed");
  LF_CHECK(op->precision == LFRIC_R64, LFRIC_E_UNSUPPORTED, "assemble_mass_w0: FP64 operator (convert for FP32)");
  LF_CHECK(nqp >= 1 && nqp <= 4, LFRIC_E_ARG, "assemble_mass_w0: 1 to 4 quadrature points per direction");
  const std::vector<double> t = w0_tables(nqp);
  void *dt = nullptr;
  if (int rc = dmalloc(&dt, sizeof(double) * t.size())) return rc;
  LF_HIP(hipMemcpyAsync(dt, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, cur_stream()));
  const long long nelem = static_cast<long long>(f.ncells) * f.nlayers;
  double *A = static_cast<double *>(op->d);
  const double *tb = static_cast<const double *>(dt);
#define ASM_CASE(N)                                                                                             \
  case N:                                                                                                       \
    LAUNCH(k_assemble_mass_w0<N>, grid_for(nelem), A, op->gs, f.d_ebase, f.d_map, chi1, chi2, chi3, tb, f.nlayers, \
           nelem, op->row_top);                                                                                 \
    break;
  switch (nqp) {
    ASM_CASE(1)
    ASM_CASE(2)
    ASM_CASE(3)
    ASM_CASE(4)
  }
#undef ASM_CASE
  const hipError_t e = hipGetLastError();
  const hipError_t es = hipStreamSynchronize(cur_stream());
  dfree(dt, sizeof(double) * t.size());
  if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
  if (es != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(es));
  return LFRIC_OK;
}

// Tables of k_assemble_fem: coordinate-basis derivatives [3][8][NP] and weights [nh], [nq] as
// in w0_tables, then the W2 basis values [6][NP]: sign * (1 - t) on W, S, B and sign * t on E,
// N, T, with t the dof's coordinate and sign that of its reference normal (-1 for S and N).
static std::vector<double> fem_tables(int nq) {
  const std::vector<double> t0 = w0_tables(nq);
  std::vector<double> p(nq), w(nq);
  gauss_rule(nq, p.data(), w.data());
  const int nh = nq * nq, np = nh * nq;
  std::vector<double> t(static_cast<size_t>(30 * np + nh + nq));
  std::copy(t0.begin() + 8 * np, t0.end(), t.begin());  // derivatives, wh, wv
  static const int comp[6] = {0, 1, 0, 1, 2, 2}, far[6] = {0, 0, 1, 1, 0, 1};
  static const double sign[6] = {1.0, -1.0, 1.0, -1.0, 1.0, 1.0};
  double *w2 = t.data() + 24 * np + nh + nq;
  for (int kz = 0; kz < nq; ++kz)
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j) {
        const int q = kz * nh + i * nq + j;
        const double xyz[3] = {p[i], p[j], p[kz]};
        for (int d = 0; d < 6; ++d) {
          const double v = xyz[comp[d]];
          w2[static_cast<size_t>(d) * np + q] = sign[d] * (far[d] ? v : 1.0 - v);
        }
      }
  return t;
}

// Roles of a lowest-order space from its dofmap: W3 has one dof, W2 six with only the top face
// (dof 6) on the next level's bottom (dof 5).
static bool is_w3_space(const FunctionSpace &f) { return f.ndf == 1 && f.top_mask == 0u; }
static bool is_w2_space(const FunctionSpace &f) {
  return f.ndf == 6 && f.top_mask == (1u << 5) && f.partner[4] == 5;
}
static bool is_wtheta_space(const FunctionSpace &f) { return f.ndf == 2 && f.top_mask == 2u && f.partner[0] == 1; }
static bool is_w0_space(const FunctionSpace &f) {
  bool w0 = f.ndf == 8 && f.top_mask == 0xF0u;
  for (int i = 0; w0 && i < 4; ++i) w0 = f.partner[i] == i + 4;
  return w0;
}

int launch_assemble_fem(lfric_op *op, int kind, const FunctionSpace &chi_fs, const double *chi1, const double *chi2,
                        const double *chi3, int nqp, int rehab) {
  const FunctionSpace &f = *op->to;
  LF_CHECK(op->precision == LFRIC_R64, LFRIC_E_UNSUPPORTED, "assemble: FP64 operator (convert for FP32)");
  LF_CHECK(nqp >= 1 && nqp <= 4, LFRIC_E_ARG, "assemble: 1 to 4 quadrature points per direction");
  LF_CHECK(is_w0_space(chi_fs), LFRIC_E_ARG, "assemble: the coordinate fields must be on a lowest-order W0 space");
  LF_CHECK(chi_fs.ncells == f.ncells && chi_fs.nlayers == f.nlayers && chi_fs.cell_of_pos == f.cell_of_pos,
           LFRIC_E_ARG, "assemble: the coordinate space is on another mesh or colouring");
  if (kind == kFemMassW3)
    LF_CHECK(op->to == op->from && is_w3_space(f), LFRIC_E_UNSUPPORTED, "assemble: a square lowest-order W3 operator is needed");
  else if (kind == kFemMassW2)
    LF_CHECK(op->to == op->from && is_w2_space(f), LFRIC_E_UNSUPPORTED, "assemble: a square lowest-order W2 operator is needed");
  else if (kind == kFemMassWtheta)
    LF_CHECK(op->to == op->from && is_wtheta_space(f), LFRIC_E_UNSUPPORTED,
             "assemble: a square lowest-order Wtheta operator is needed");
  else
    LF_CHECK(is_w3_space(f) && is_w2_space(*op->from) && op->from->cell_of_pos == f.cell_of_pos, LFRIC_E_UNSUPPORTED,
             "assemble: the divergence maps a lowest-order W2 space to a W3 space of the same mesh");
  const std::vector<double> t = fem_tables(nqp);
  void *dt = nullptr;
  if (int rc = dmalloc(&dt, sizeof(double) * t.size())) return rc;
  LF_HIP(hipMemcpyAsync(dt, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, cur_stream()));
  const long long nelem = static_cast<long long>(f.ncells) * f.nlayers;
  double *A = static_cast<double *>(op->d);
  const double *tb = static_cast<const double *>(dt);
#define FEM_CASE(N, K)                                                                                           \
  case N * 4 + K:                                                                                                \
    LAUNCH((k_assemble_fem<N, K>), grid_for(nelem), A, op->gs, f.d_ebase, chi_fs.d_map, chi1, chi2, chi3, tb,      \
           f.nlayers, nelem, op->row_top, rehab);                                                               \
    break;
  switch (nqp * 4 + kind) {
    FEM_CASE(1, 0) FEM_CASE(1, 1) FEM_CASE(1, 2) FEM_CASE(1, 3)
    FEM_CASE(2, 0) FEM_CASE(2, 1) FEM_CASE(2, 2) FEM_CASE(2, 3)
    FEM_CASE(3, 0) FEM_CASE(3, 1) FEM_CASE(3, 2) FEM_CASE(3, 3)
    FEM_CASE(4, 0) FEM_CASE(4, 1) FEM_CASE(4, 2) FEM_CASE(4, 3)
    default:
      dfree(dt, sizeof(double) * t.size());
      return fail(LFRIC_E_ARG, "assemble: unknown operator kind");
  }
#undef FEM_CASE
  const hipError_t e = hipGetLastError();
  const hipError_t es = hipStreamSynchronize(cur_stream());
  dfree(dt, sizeof(double) * t.size());
  if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
  if (es != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(es));
  return LFRIC_OK;
}

int launch_op_invert_local(const lfric_op *A, lfric_op *Ainv) {
  LF_CHECK(A->to == A->from, LFRIC_E_ARG, "invert_local: the operator must map a space to itself");
  LF_CHECK(Ainv->to == A->to && Ainv->from == A->from, LFRIC_E_ARG, "invert_local: output on other spaces");
  LF_CHECK(Ainv->precision == A->precision, LFRIC_E_ARG, "invert_local: output of another precision");
  LF_CHECK(A->to->ndf <= 32, LFRIC_E_UNSUPPORTED, "invert_local: more than 32 dofs per cell");
  if (A->precision == LFRIC_R64) return invert_local_t<double>(A, Ainv);
  return invert_local_t<float>(A, Ainv);
}

int launch_convert(int pd, void *dst, int ps, const void *src, long long n, double scale) {
  const int grid = grid_for(n, kBlock, 65536);
  if (pd == LFRIC_R64 && ps == LFRIC_R64)
    LAUNCH((k_convert<double, double>), grid, static_cast<double *>(dst), static_cast<const double *>(src), n, scale);
  else if (pd == LFRIC_R32 && ps == LFRIC_R64)
    LAUNCH((k_convert<float, double>), grid, static_cast<float *>(dst), static_cast<const double *>(src), n, scale);
  else if (pd == LFRIC_R64 && ps == LFRIC_R32)
    LAUNCH((k_convert<double, float>), grid, static_cast<double *>(dst), static_cast<const float *>(src), n, scale);
  else if (pd == LFRIC_R32 && ps == LFRIC_R32)
    LAUNCH((k_convert<float, float>), grid, static_cast<float *>(dst), static_cast<const float *>(src), n, scale);
  else
    return fail(LFRIC_E_UNSUPPORTED, "convert: unsupported precision pair");
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_op_convert(const lfric_op *src, lfric_op *dst) {
  LF_CHECK(src->ngroups == dst->ngroups, LFRIC_E_ARG, "operator shapes differ");
  const long long nelem = static_cast<long long>(src->ngroups) * 64;
  const int np = src->to->ndf * src->from->ndf;
  const int grid = grid_for(nelem * np);
  const bool d64 = dst->precision == LFRIC_R64, s64 = src->precision == LFRIC_R64;
  if (d64 && s64)
    LAUNCH((k_op_convert<double, double>), grid, static_cast<double *>(dst->d), dst->gs,
           static_cast<const double *>(src->d), src->gs, nelem, np);
  else if (!d64 && s64)
    LAUNCH((k_op_convert<float, double>), grid, static_cast<float *>(dst->d), dst->gs,
           static_cast<const double *>(src->d), src->gs, nelem, np);
  else if (d64 && !s64)
    LAUNCH((k_op_convert<double, float>), grid, static_cast<double *>(dst->d), dst->gs,
           static_cast<const float *>(src->d), src->gs, nelem, np);
  else
    LAUNCH((k_op_convert<float, float>), grid, static_cast<float *>(dst->d), dst->gs,
           static_cast<const float *>(src->d), src->gs, nelem, np);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

// ---- operator apply dispatch ---------------------------------------------------------------
// which compile-time role table (if any) a space's detected roles match
enum class RoleKind { None, W0K0, W0K1, W2K0, Runtime };
template <class R>
static bool roles_match(const FunctionSpace &f) {
  for (int i = 0; i < f.ndf; ++i) {
    if (R::top(i) != static_cast<bool>((f.top_mask >> i) & 1u)) return false;
    if (R::partner(i) != f.partner[i]) return false;
  }
  return true;
}
static RoleKind role_kind(const FunctionSpace &f) {
  if (f.ndf == 8 && roles_match<RolesW0K0>(f)) return RoleKind::W0K0;
  if (f.ndf == 27 && roles_match<RolesW0K1>(f)) return RoleKind::W0K1;
  if (f.ndf == 6 && roles_match<RolesW2K0>(f)) return RoleKind::W2K0;
  if (roles_match<RolesNone>(f)) return RoleKind::None;
  return RoleKind::Runtime;
}

template <typename T, int A_, int B_>
static void launch_colour(RoleKind rk, int grid, const lfric_op *op, size_t S, const T *x, T *y, const FunctionSpace &to,
                          const FunctionSpace &from, const uint32_t *first, int pos0, int npos, const int *d_skip) {
  const int nl = to.nlayers;
  const T *Ad = static_cast<const T *>(op->d);
  switch (rk) {
    case RoleKind::W0K0:
      if constexpr (A_ == 8)
        LAUNCH((k_lma_colour<T, A_, B_, RolesW0K0>), grid, Ad, S, to.d_ebase, x, y, to.d_map, from.d_map, first, pos0, npos, nl,
               RolesW0K0{}, d_skip, op->xwin);
      break;
    case RoleKind::W0K1:
      if constexpr (A_ == 27)
        LAUNCH((k_lma_colour<T, A_, B_, RolesW0K1>), grid, Ad, S, to.d_ebase, x, y, to.d_map, from.d_map, first, pos0, npos, nl,
               RolesW0K1{}, d_skip, op->xwin);
      break;
    case RoleKind::W2K0:
      if constexpr (A_ == 6)
        LAUNCH((k_lma_colour<T, A_, B_, RolesW2K0>), grid, Ad, S, to.d_ebase, x, y, to.d_map, from.d_map, first, pos0, npos, nl,
               RolesW2K0{}, d_skip, op->xwin);
      break;
    case RoleKind::None:
      LAUNCH((k_lma_colour<T, A_, B_, RolesNone>), grid, Ad, S, to.d_ebase, x, y, to.d_map, from.d_map, first, pos0, npos, nl,
             RolesNone{}, d_skip, op->xwin);
      break;
    default:
      LAUNCH((k_lma_colour<T, A_, B_, RolesRuntime>), grid, Ad, S, to.d_ebase, x, y, to.d_map, from.d_map, first, pos0, npos, nl,
             RolesRuntime{roles_of(to)}, d_skip, op->xwin);
  }
}

template <typename T>
static int apply_colours_t(const lfric_op *op, const T *x, T *y, const int *d_skip) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  RoleKind rk = role_kind(to);
  const size_t S = op->gs;
  const int nl = to.nlayers;
  for (int c = 0; c < to.ncolours; ++c) {
    const int pos0 = to.colour_start[c], npos = to.colour_start[c + 1] - pos0;
    if (npos == 0) continue;
    const int grid = grid_for(static_cast<long long>(npos) * (nl + 1));
#define LMA_CASE(A, B)                                                                              \
  if (to.ndf == A && from.ndf == B) {                                                              \
    launch_colour<T, A, B>(rk, grid, op, S, x, y, to, from, to.d_first, pos0, npos, d_skip);    \
    continue;                                                                                       \
  }
    LMA_CASE(8, 8)
    LMA_CASE(27, 27)
    LMA_CASE(6, 6)
    LMA_CASE(1, 1)
    LMA_CASE(1, 6)
    LMA_CASE(6, 1)
    LMA_CASE(1, 8)
    LMA_CASE(8, 1)
    LMA_CASE(4, 2)
#undef LMA_CASE
    return fail(LFRIC_E_UNSUPPORTED, "lma_apply: unsupported (ndf_to, ndf_from)");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_lma_apply(const lfric_op *op, const void *x, void *y, const int *d_skip) {
  if (op->precision == LFRIC_R64)
    return apply_colours_t<double>(op, static_cast<const double *>(x), static_cast<double *>(y), d_skip);
  if (op->precision == LFRIC_R32)
    return apply_colours_t<float>(op, static_cast<const float *>(x), static_cast<float *>(y), d_skip);
  return fail(LFRIC_E_UNSUPPORTED, "lma_apply: operator precision");
}

// column lists of the gather form (host build, once per space)
static int build_gather(FunctionSpace &f) {
  if (f.d_gcol) return LFRIC_OK;
  const int nc = f.ncells, ndf = f.ndf;
  LF_CHECK(static_cast<long long>(nc) * 32 < (1LL << 31), LFRIC_E_UNSUPPORTED, "mesh too large for the gather form");
  struct Item {
    int base, pos, df;
  };
  std::vector<Item> items;
  items.reserve(static_cast<size_t>(nc) * ndf);
  for (int p = 0; p < nc; ++p) {
    const int cell = f.cell_of_pos[p];
    for (int d = 0; d < ndf; ++d)
      if (!((f.top_mask >> d) & 1u)) items.push_back({f.dofmap[static_cast<size_t>(cell) * ndf + d] - 1, p, d});
  }
  std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.base < b.base; });
  std::vector<int2> cols;
  std::vector<int> off, ent;
  ent.reserve(items.size());
  for (size_t q = 0; q < items.size(); ++q) {
    if (q == 0 || items[q].base != items[q - 1].base) {
      cols.push_back(make_int2(items[q].base, f.partner[items[q].df] >= 0 ? 1 : 0));
      off.push_back(static_cast<int>(ent.size()));
    } else {
      LF_CHECK((f.partner[items[q].df] >= 0) == (cols.back().y == 1), LFRIC_E_UNSUPPORTED,
               "cells disagree on the height of a dof column");
    }
    ent.push_back(items[q].pos * 32 + items[q].df);
  }
  off.push_back(static_cast<int>(ent.size()));
  f.gather_ncols = static_cast<int>(cols.size());
  f.gcol_bytes = sizeof(int2) * cols.size();
  f.goff_bytes = sizeof(int) * off.size();
  f.gent_bytes = sizeof(int) * ent.size();
  if (int rc = dmalloc(reinterpret_cast<void **>(&f.d_gcol), f.gcol_bytes)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&f.d_goff), f.goff_bytes)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&f.d_gent), f.gent_bytes)) return rc;
  LF_HIP(hipMemcpy(f.d_gcol, cols.data(), f.gcol_bytes, hipMemcpyHostToDevice));
  LF_HIP(hipMemcpy(f.d_goff, off.data(), f.goff_bytes, hipMemcpyHostToDevice));
  LF_HIP(hipMemcpy(f.d_gent, ent.data(), f.gent_bytes, hipMemcpyHostToDevice));
  return LFRIC_OK;
}

int gather_blocks(const lfric_op *op) {
  const FunctionSpace &to = *op->to;
  return grid_for(static_cast<long long>(to.gather_ncols) * (to.nlayers + 1));
}

int gather_dot_blocks(const lfric_op *op) {
  const bool sym = op->sym && op->to == op->from && role_kind(*op->to) == RoleKind::W0K0;
  return sym ? kRedBlocks : gather_blocks(op);
}

// set by lfric_op_time_apply: an event recorded between the two phases of the symmetric-packed
// apply, so that each phase can be timed on its own
static hipEvent_t g_phase_mark = nullptr;

int launch_lma_apply_gather_marked(const lfric_op *op, const void *x, void *y, hipEvent_t mark, bool *two_phase) {
  *two_phase = op->sym != nullptr && op->to == op->from &&
               (role_kind(*op->to) == RoleKind::W0K0 || role_kind(*op->to) == RoleKind::W0K1);
  g_phase_mark = *two_phase ? mark : nullptr;
  const int rc = launch_lma_apply_gather(op, x, y);
  g_phase_mark = nullptr;
  return rc;
}

struct GatherExtra {
  double *partials = nullptr;  // non-null: fused x.y partial sums, one per block
  const int *skip = nullptr;
};

// the two phases of the order-1 symmetric-packed apply (sym27_apply.h), NR right-hand sides;
// with partials (one RHS) the second phase adds the x.y partial sum of each of its blocks
template <typename T, int NR>
static int launch_sym27(const lfric_op *op, const T *x, T *y, double *partials, const int *skip) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  const long long units = static_cast<long long>(to.ncells) * to.nlayers;
  constexpr int NRL = s27_nrl<T, NR>();
  T *z = static_cast<T *>(op->symz);
  if (NR > 1) {
    const size_t need = sizeof(T) * s27_zalloc(to.ncells, to.nlayers, NR);
    if (op->symz_mr_bytes < need) {
      dfree(op->symz_mr, op->symz_mr_bytes);
      op->symz_mr = nullptr;
      op->symz_mr_bytes = 0;
      if (int rc = dmalloc(&op->symz_mr, need)) return rc;
      op->symz_mr_bytes = need;
    }
    z = static_cast<T *>(op->symz_mr);
  }
  LAUNCH((k_sym27_units<T, NR, NRL>), grid_for(s27_threads<T, NR>(units)), static_cast<const T *>(op->sym), op->gss,
         x, from.d_map, op->d_sympos, z, units, to.nlayers, skip, op->xwin);
  if (g_phase_mark) (void)hipEventRecord(g_phase_mark, cur_stream());
  const int grid = gather_blocks(op);
  if (partials) {
    if constexpr (NR == 1)
      LAUNCH((k_sym27_gather<T, 1, true>), grid, z, x, y, to.d_gcol, op->d_symzu, to.gather_ncols, to.nlayers,
             partials, skip);
  } else {
    LAUNCH((k_sym27_gather<T, NR, false>), grid, z, x, y, to.d_gcol, op->d_symzu, to.gather_ncols, to.nlayers,
           partials, skip);
  }
  return LFRIC_OK;
}

template <typename T, int N2, class R>
static void launch_gather_r(const lfric_op *op, const T *x, T *y, R roles, GatherExtra ex) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  const int grid = gather_blocks(op);
  if constexpr (N2 == 27 && std::is_same<R, RolesW0K1>::value) {
    if (sym27_ready(op)) {
      (void)launch_sym27<T, 1>(op, x, y, ex.partials, ex.skip);
      return;
    }
  }
  if constexpr (N2 == 8 && std::is_same<R, RolesW0K0>::value) {
    if (op->sym) {  // symmetric-packed copy (launch_build_sym): two phases
      const int nseg = op->sym_seg > 0 ? (to.nlayers + 1 + op->sym_seg - 1) / op->sym_seg : 0;
      if (op->sym_seg == 0)
        LAUNCH((k_sym_units1<T>), grid_for(s1_threads(static_cast<long long>(to.ncells) * (to.nlayers + 1))),
               static_cast<const T *>(op->sym), op->gss, x, op->d_symcol, static_cast<T *>(op->symz), to.ncells,
               to.nlayers, ex.skip, op->xwin);
      else if (op->sym_seg == 4)
        LAUNCH((k_sym_units_s<T, 4>), grid_for(static_cast<long long>(to.ncells) * nseg),
               static_cast<const T *>(op->sym), op->gss, x, from.d_map, static_cast<T *>(op->symz), to.ncells,
               to.nlayers, nseg, ex.skip);
      else if (op->sym_seg == 2)
        LAUNCH((k_sym_units_s<T, 2>), grid_for(static_cast<long long>(to.ncells) * nseg),
               static_cast<const T *>(op->sym), op->gss, x, from.d_map, static_cast<T *>(op->symz), to.ncells,
               to.nlayers, nseg, ex.skip);
      else
        LAUNCH((k_sym_units<T>), grid_for(static_cast<long long>(to.ncells) * nseg), static_cast<const T *>(op->sym),
               op->gss, x, from.d_map, static_cast<T *>(op->symz), to.ncells, to.nlayers, op->sym_seg, nseg, ex.skip);
      if (g_phase_mark) (void)hipEventRecord(g_phase_mark, cur_stream());
      if (ex.partials)
        LAUNCH((k_sym_gather_dot<T>), kRedBlocks, static_cast<const T *>(op->symz), x, y, to.d_gcol, to.d_goff,
               to.d_gent, op->d_symzu, to.gather_ncols, to.nlayers, ex.partials, ex.skip, op->xwin);
      else
        LAUNCH((k_sym_gather<T>), grid, static_cast<const T *>(op->symz), x, y, to.d_gcol, to.d_goff, to.d_gent,
               op->d_symzu, to.gather_ncols, to.nlayers, ex.skip, op->xwin);
      return;
    }
    if (op->rows) {  // built only for square lowest-order W0 operators (launch_build_rows)
      const T *rw = static_cast<const T *>(op->rows);
      if (ex.partials)
        LAUNCH((k_lma_gather_rows_w0<T, true>), grid, rw, op->gsr, x, y, from.d_map, to.d_gcol, to.d_goff,
               to.d_gent, to.gather_ncols, to.nlayers, ex.partials, ex.skip, op->xwin);
      else
        LAUNCH((k_lma_gather_rows_w0<T, false>), grid, rw, op->gsr, x, y, from.d_map, to.d_gcol, to.d_goff,
               to.d_gent, to.gather_ncols, to.nlayers, ex.partials, ex.skip, op->xwin);
      return;
    }
  }
  if (ex.partials)
    LAUNCH((k_lma_gather<T, N2, R, true>), grid, static_cast<const T *>(op->d), op->gs, to.d_ebase, x, y, from.d_map,
           to.d_gcol, to.d_goff, to.d_gent, to.gather_ncols, to.nlayers, roles, ex.partials, ex.skip, op->xwin);
  else
    LAUNCH((k_lma_gather<T, N2, R, false>), grid, static_cast<const T *>(op->d), op->gs, to.d_ebase, x, y,
           from.d_map, to.d_gcol, to.d_goff, to.d_gent, to.gather_ncols, to.nlayers, roles, ex.partials, ex.skip, op->xwin);
}

template <typename T, int N2>
static void launch_gather_n(RoleKind rk, const lfric_op *op, const T *x, T *y, GatherExtra ex) {
  switch (rk) {
    case RoleKind::W0K0: launch_gather_r<T, N2>(op, x, y, RolesW0K0{}, ex); break;
    case RoleKind::W0K1: launch_gather_r<T, N2>(op, x, y, RolesW0K1{}, ex); break;
    case RoleKind::W2K0: launch_gather_r<T, N2>(op, x, y, RolesW2K0{}, ex); break;
    case RoleKind::None: launch_gather_r<T, N2>(op, x, y, RolesNone{}, ex); break;
    default: launch_gather_r<T, N2>(op, x, y, RolesRuntime{roles_of(*op->to)}, ex);
  }
}

template <typename T>
static int gather_t(const lfric_op *op, const T *x, T *y, GatherExtra ex) {
  if (int rc = build_gather(*op->to)) return rc;
  const RoleKind rk = role_kind(*op->to);
  switch (op->from->ndf) {
    case 1: launch_gather_n<T, 1>(rk, op, x, y, ex); break;
    case 2: launch_gather_n<T, 2>(rk, op, x, y, ex); break;
    case 6: launch_gather_n<T, 6>(rk, op, x, y, ex); break;
    case 8: launch_gather_n<T, 8>(rk, op, x, y, ex); break;
    case 27: launch_gather_n<T, 27>(rk, op, x, y, ex); break;
    default: return fail(LFRIC_E_UNSUPPORTED, "lma_apply_gather: unsupported ndf_from");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

static int gather_any(const lfric_op *op, const void *x, void *y, GatherExtra ex) {
  if (op->precision == LFRIC_R64)
    return gather_t<double>(op, static_cast<const double *>(x), static_cast<double *>(y), ex);
  if (op->precision == LFRIC_R32)
    return gather_t<float>(op, static_cast<const float *>(x), static_cast<float *>(y), ex);
  return fail(LFRIC_E_UNSUPPORTED, "lma_apply_gather: operator precision");
}

// ---- multi-RHS dispatch ---------------------------------------------------------------------
template <typename T, int N2, int NR, class R>
static int launch_gather_mrhs_r(const lfric_op *op, const T *x, T *y, R roles) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  const int grid = grid_for(static_cast<long long>(to.gather_ncols) * (to.nlayers + 1));
  if constexpr (N2 == 27 && std::is_same<R, RolesW0K1>::value) {
    if (sym27_ready(op)) return launch_sym27<T, NR>(op, x, y, nullptr, nullptr);
  }
  if constexpr (N2 == 8 && std::is_same<R, RolesW0K0>::value) {
    if (op->rows) {  // the row-gathered copy (lowest-order W0, launch_build_rows)
      LAUNCH((k_lma_gather_rows_w0_mrhs<T, NR>), grid, static_cast<const T *>(op->rows), op->gsr, x, y, from.d_map,
             to.d_gcol, to.d_goff, to.d_gent, to.gather_ncols, to.nlayers, op->xwin);
      return LFRIC_OK;
    }
  }
  LAUNCH((k_lma_gather_mrhs<T, N2, NR, R>), grid, static_cast<const T *>(op->d), op->gs, to.d_ebase, x, y,
         from.d_map, to.d_gcol, to.d_goff, to.d_gent, to.gather_ncols, to.nlayers, roles, op->xwin);
  return LFRIC_OK;
}

template <typename T, int N1, int N2, int NR, class R>
static void launch_colour_mrhs_r(const lfric_op *op, const T *x, T *y, R roles) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  for (int c = 0; c < to.ncolours; ++c) {
    const int pos0 = to.colour_start[c], npos = to.colour_start[c + 1] - pos0;
    if (npos == 0) continue;
    LAUNCH((k_lma_colour_mrhs<T, N1, N2, NR, R>), grid_for(static_cast<long long>(npos) * (to.nlayers + 1)),
           static_cast<const T *>(op->d), op->gs, to.d_ebase, x, y, to.d_map, from.d_map, to.d_first, pos0, npos,
           to.nlayers, roles, op->xwin);
  }
}

template <typename T, int N1, int N2, int NR>
static int launch_mrhs_n(int variant, RoleKind rk, const lfric_op *op, const T *x, T *y) {
  const Roles rr = roles_of(*op->to);
  if (variant == LFRIC_MRHS_COLOUR) {
    switch (rk) {
      case RoleKind::W0K0: launch_colour_mrhs_r<T, N1, N2, NR>(op, x, y, RolesW0K0{}); break;
      case RoleKind::W0K1: launch_colour_mrhs_r<T, N1, N2, NR>(op, x, y, RolesW0K1{}); break;
      case RoleKind::W2K0: launch_colour_mrhs_r<T, N1, N2, NR>(op, x, y, RolesW2K0{}); break;
      case RoleKind::None: launch_colour_mrhs_r<T, N1, N2, NR>(op, x, y, RolesNone{}); break;
      default: launch_colour_mrhs_r<T, N1, N2, NR>(op, x, y, RolesRuntime{rr});
    }
    return LFRIC_OK;
  }
  switch (rk) {
    case RoleKind::W0K0: return launch_gather_mrhs_r<T, N2, NR>(op, x, y, RolesW0K0{});
    case RoleKind::W0K1: return launch_gather_mrhs_r<T, N2, NR>(op, x, y, RolesW0K1{});
    case RoleKind::W2K0: return launch_gather_mrhs_r<T, N2, NR>(op, x, y, RolesW2K0{});
    case RoleKind::None: return launch_gather_mrhs_r<T, N2, NR>(op, x, y, RolesNone{});
    default: return launch_gather_mrhs_r<T, N2, NR>(op, x, y, RolesRuntime{rr});
  }
}

template <typename T, int NR>
static int mrhs_nr(int variant, const lfric_op *op, const T *x, T *y) {
  const RoleKind rk = role_kind(*op->to);
  const int n1 = op->to->ndf, n2 = op->from->ndf;
#define MRHS_CASE(A, B)                                      \
  if (n1 == A && n2 == B) {                                  \
    if (int rc = launch_mrhs_n<T, A, B, NR>(variant, rk, op, x, y)) return rc; \
    CHECK_LAUNCH();                                          \
    return LFRIC_OK;                                         \
  }
  MRHS_CASE(8, 8)
  MRHS_CASE(27, 27)
  MRHS_CASE(6, 6)
  MRHS_CASE(1, 1)
#undef MRHS_CASE
  return fail(LFRIC_E_UNSUPPORTED, "lma_apply_mrhs: unsupported (ndf_to, ndf_from)");
}

template <typename T>
static int mrhs_t(int variant, int nrhs, const lfric_op *op, const T *x, T *y) {
  if (variant == LFRIC_MRHS_GATHER)
    if (int rc = build_gather(*op->to)) return rc;
  switch (nrhs) {
    case 2: return mrhs_nr<T, 2>(variant, op, x, y);
    case 4: return mrhs_nr<T, 4>(variant, op, x, y);
    case 8: return mrhs_nr<T, 8>(variant, op, x, y);
  }
  return fail(LFRIC_E_UNSUPPORTED, "lma_apply_mrhs: nrhs must be 1, 2, 4 or 8");
}

template <int NR>
static void launch_mrhs_mfma_nr(const lfric_op *op, const float *x, float *y) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  for (int c = 0; c < to.ncolours; ++c) {
    const int pos0 = to.colour_start[c], npos = to.colour_start[c + 1] - pos0;
    if (npos == 0) continue;
    LAUNCH((k_lma_colour_mrhs_mfma<NR>), grid_for(static_cast<long long>(npos) * (to.nlayers + 1), kBlock / 4),
           static_cast<const float *>(op->d), op->gs, to.d_ebase, x, y, to.d_map, from.d_map, to.d_first, pos0, npos,
           to.nlayers, op->xwin);
  }
}

bool mfma_apply_supported(const lfric_op *op) {
  return op->precision == LFRIC_R32 && op->to == op->from && op->to->ndf == 8 && role_kind(*op->to) == RoleKind::W0K0;
}

int batch_colour_form(const lfric_op *op) {
  if (mfma_apply_supported(op)) return 1;
  if (op->precision == LFRIC_R32 && op->to == op->from && op->to->ndf == 27 && role_kind(*op->to) == RoleKind::W0K1)
    return 2;
  return 0;
}

static int mrhs_mfma(const lfric_op *op, const void *xv, void *yv, int nrhs) {
  LF_CHECK(mfma_apply_supported(op), LFRIC_E_UNSUPPORTED,
           "lma_apply_mrhs: the MFMA variant is for square lowest-order W0 FP32 operators");
  const float *x = static_cast<const float *>(xv);
  float *y = static_cast<float *>(yv);
  switch (nrhs) {
    case 1: launch_mrhs_mfma_nr<1>(op, x, y); break;
    case 2: launch_mrhs_mfma_nr<2>(op, x, y); break;
    case 4: launch_mrhs_mfma_nr<4>(op, x, y); break;
    case 8: launch_mrhs_mfma_nr<8>(op, x, y); break;
    default: return fail(LFRIC_E_UNSUPPORTED, "lma_apply_mrhs: nrhs must be 1, 2, 4 or 8");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_lma_apply_mrhs(const lfric_op *op, const void *x, void *y, int nrhs, int variant) {
  if (variant == LFRIC_MRHS_AUTO) {
    // measured: an operator with a row-gathered copy is fastest in the gather form over the
    // copy; otherwise the colour-ordered form with register-staged x is the fastest from a few
    // hundred thousand cells up (C96x70 W0, 4 FP32 RHS: 0.310 ms vs gather 0.354); on small
    // meshes its one launch per colour leaves the GPU idle and the gather form wins
    const long long nc3 = static_cast<long long>(op->to->ncells) * op->to->nlayers;
    variant = (nrhs > 1 && nc3 >= 200000 && !op->rows && !sym27_ready(op)) ? LFRIC_MRHS_COLOUR : LFRIC_MRHS_GATHER;
  }
  LF_CHECK(variant == LFRIC_MRHS_GATHER || variant == LFRIC_MRHS_COLOUR || variant == LFRIC_MRHS_MFMA, LFRIC_E_ARG,
           "unknown multi-RHS variant");
  if (variant == LFRIC_MRHS_MFMA) return mrhs_mfma(op, x, y, nrhs);
  if (nrhs == 1)
    return variant == LFRIC_MRHS_GATHER ? launch_lma_apply_gather(op, x, y) : launch_lma_apply(op, x, y, nullptr);
  if (op->precision == LFRIC_R64)
    return mrhs_t<double>(variant, nrhs, op, static_cast<const double *>(x), static_cast<double *>(y));
  return mrhs_t<float>(variant, nrhs, op, static_cast<const float *>(x), static_cast<float *>(y));
}

int launch_lma_apply_gather(const lfric_op *op, const void *x, void *y) { return gather_any(op, x, y, GatherExtra{}); }

// build the row-gathered copy of a square operator (see k_build_rows)
template <typename T, int N2, class R>
static int build_rows_r(lfric_op *op, R roles) {
  const FunctionSpace &to = *op->to;
  constexpr int RS = rows_rs<T, N2>();
  const long long threads = static_cast<long long>(to.gather_ncols) * (to.nlayers + 1);
  const size_t gsr = static_cast<size_t>(64) * 2 * kRowsKmax * RS;
  const size_t bytes = sizeof(T) * gsr * static_cast<size_t>((threads + 63) / 64);
  void *rows = nullptr;
  if (int rc = dmalloc(&rows, bytes)) return rc;
  LAUNCH((k_build_rows<T, N2, R>), grid_for(threads), static_cast<const T *>(op->d), op->gs, to.d_ebase,
         static_cast<T *>(rows), gsr, to.d_gcol, to.d_goff, to.d_gent, to.gather_ncols, to.nlayers, roles);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    dfree(rows, bytes);
    return fail(LFRIC_E_HIP, hipGetErrorString(e));
  }
  dfree(op->rows, op->rows_bytes);
  op->rows = rows;
  op->rows_bytes = bytes;
  op->gsr = gsr;
  return LFRIC_OK;
}

int launch_build_rows(lfric_op *op) {
  FunctionSpace &to = *op->to;
  LF_CHECK(op->to == op->from, LFRIC_E_ARG, "the row-gathered copy is for square operators");
  if (int rc = build_gather(to)) return rc;
  std::vector<int> off(static_cast<size_t>(to.gather_ncols) + 1);
  LF_HIP(hipMemcpy(off.data(), to.d_goff, sizeof(int) * off.size(), hipMemcpyDeviceToHost));
  for (int c = 0; c < to.gather_ncols; ++c)
    LF_CHECK(off[c + 1] - off[c] <= kRowsKmax, LFRIC_E_UNSUPPORTED, "a dof column takes rows from more than 4 cells");
  const RoleKind rk = role_kind(to);
  LF_CHECK(rk == RoleKind::W0K0 && to.ndf == 8, LFRIC_E_UNSUPPORTED,
           "the row-gathered copy is built for lowest-order W0 operators");
  for (int c = 0; c < to.ncells; ++c)
    for (int j = 0; j < 4; ++j)
      LF_CHECK(to.dofmap[static_cast<size_t>(c) * 8 + j + 4] == to.dofmap[static_cast<size_t>(c) * 8 + j] + 1,
               LFRIC_E_UNSUPPORTED, "the row-gathered copy needs each top vertex one level above its bottom vertex");
  return op->precision == LFRIC_R64 ? build_rows_r<double, 8>(op, RolesW0K0{}) : build_rows_r<float, 8>(op, RolesW0K0{});
}

// device buffers of a symmetric-packed copy under construction, freed unless handed over
struct SymBuild {
  void *sym = nullptr, *z = nullptr;
  int4 *zu = nullptr, *col = nullptr;
  int *slotpos = nullptr, *flag = nullptr;
  size_t sym_b = 0, z_b = 0, zu_b = 0, col_b = 0, slotpos_b = 0;
  ~SymBuild() {
    dfree(sym, sym_b);
    dfree(z, z_b);
    dfree(zu, zu_b);
    dfree(col, col_b);
    dfree(slotpos, slotpos_b);
    dfree(flag, sizeof(int));
  }
};

template <typename T>
static int build_sym_t(lfric_op *op, int S) {
  const FunctionSpace &f = *op->to;
  constexpr int V = chunk_entries<T>(), NCH = 36 / V;
  const int nc = f.ncells, H = f.nlayers + 1;
  // S = 0: one unit per thread, one triangle per slot, cells in cell-id order; else S+1
  // triangles per thread, cells in colour-major position order
  const int nseg = S > 0 ? (H + S - 1) / S : H;
  const long long threads = static_cast<long long>(nc) * nseg;
  const size_t gss = static_cast<size_t>(64) * (S > 0 ? S + 1 : 1) * NCH * V;
  SymBuild b;
  b.sym_b = sizeof(T) * gss * static_cast<size_t>((threads + 63) / 64);
  b.z_b = sizeof(T) * static_cast<size_t>(nc) * 4 * H;
  if (int rc = dmalloc(&b.sym, b.sym_b)) return rc;
  if (int rc = dmalloc(&b.z, b.z_b)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&b.flag), sizeof(int))) return rc;
  // unit id of (position, bottom vertex i): slot-cell * 4 + i
  std::vector<int> unit_cell(nc);
  for (int p = 0; p < nc; ++p) unit_cell[p] = S == 0 ? f.cell_of_pos[p] : p;
  if (S == 0) {
    std::vector<int> slotpos(nc);
    std::vector<int4> col(nc);
    for (int p = 0; p < nc; ++p) slotpos[f.cell_of_pos[p]] = p;
    for (int k = 0; k < nc; ++k) {
      const int *dm = &f.dofmap[static_cast<size_t>(k) * 8];
      col[k] = make_int4(dm[0] - 1, dm[1] - 1, dm[2] - 1, dm[3] - 1);
    }
    b.slotpos_b = sizeof(int) * nc;
    b.col_b = sizeof(int4) * nc;
    if (int rc = dmalloc(reinterpret_cast<void **>(&b.slotpos), b.slotpos_b)) return rc;
    if (int rc = dmalloc(reinterpret_cast<void **>(&b.col), b.col_b)) return rc;
    LF_HIP(hipMemcpy(b.slotpos, slotpos.data(), b.slotpos_b, hipMemcpyHostToDevice));
    LF_HIP(hipMemcpy(b.col, col.data(), b.col_b, hipMemcpyHostToDevice));
  }
  // the gather's per-column unit table (columns of at most 4 cells), in colour order
  {
    const int ncol = f.gather_ncols;
    std::vector<int> off(ncol + 1), ent;
    LF_HIP(hipMemcpy(off.data(), f.d_goff, sizeof(int) * (ncol + 1), hipMemcpyDeviceToHost));
    ent.resize(off[ncol]);
    LF_HIP(hipMemcpy(ent.data(), f.d_gent, sizeof(int) * ent.size(), hipMemcpyDeviceToHost));
    bool fits = true;
    for (int c = 0; fits && c < ncol; ++c) fits = off[c + 1] - off[c] <= 4;
    if (fits) {
      std::vector<int4> h(ncol);
      for (int c = 0; c < ncol; ++c) {
        int u[4] = {-1, -1, -1, -1};
        for (int q = off[c]; q < off[c + 1]; ++q) u[q - off[c]] = unit_cell[ent[q] >> 5] * 4 + (ent[q] & 31);
        h[c] = make_int4(u[0], u[1], u[2], u[3]);
      }
      b.zu_b = sizeof(int4) * h.size();
      if (int rc = dmalloc(reinterpret_cast<void **>(&b.zu), b.zu_b)) return rc;
      LF_HIP(hipMemcpy(b.zu, h.data(), b.zu_b, hipMemcpyHostToDevice));
    } else {
      LF_CHECK(S > 0, LFRIC_E_UNSUPPORTED,
               "the one-unit-per-thread symmetric copy needs at most 4 cells around each vertex column");
    }
  }
  int h_flag = 0;
  LF_HIP(hipMemcpyAsync(b.flag, &h_flag, sizeof(int), hipMemcpyHostToDevice, cur_stream()));
  if (S == 0)
    LAUNCH((k_build_sym1<T>), grid_for(threads), static_cast<const T *>(op->d), op->gs, f.d_ebase, b.slotpos,
           static_cast<T *>(b.sym), gss, nc, f.nlayers, op->row_top, b.flag);
  else
    LAUNCH((k_build_sym<T>), grid_for(threads), static_cast<const T *>(op->d), op->gs, f.d_ebase,
           static_cast<T *>(b.sym), gss, nc, f.nlayers, S, nseg, op->row_top, b.flag);
  CHECK_LAUNCH();
  LF_HIP(hipMemcpyAsync(&h_flag, b.flag, sizeof(int), hipMemcpyDeviceToHost, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  LF_CHECK(!h_flag, LFRIC_E_UNSUPPORTED, "the operator's element matrices are not symmetric");
  drop_sym(op);
  op->sym = b.sym;
  op->sym_bytes = b.sym_b;
  op->symz = b.z;
  op->symz_bytes = b.z_b;
  op->d_symzu = b.zu;
  op->symzu_bytes = b.zu_b;
  op->d_symcol = b.col;
  op->symcol_bytes = b.col_b;
  op->gss = gss;
  op->sym_seg = S;
  b.sym = b.z = nullptr;
  b.zu = b.col = nullptr;
  return LFRIC_OK;
}

void drop_sym(lfric_op *op) {
  dfree(op->sym, op->sym_bytes);
  dfree(op->symz, op->symz_bytes);
  dfree(op->d_symzu, op->symzu_bytes);
  dfree(op->d_symcol, op->symcol_bytes);
  dfree(op->d_sympos, op->sympos_bytes);
  dfree(op->symz_mr, op->symz_mr_bytes);
  op->sym = op->symz = op->symz_mr = nullptr;
  op->d_symzu = op->d_symcol = nullptr;
  op->d_sympos = nullptr;
  op->sym_bytes = op->symz_bytes = op->symzu_bytes = op->symcol_bytes = op->sympos_bytes = op->symz_mr_bytes = 0;
  op->gss = 0;
  op->sym_seg = 0;
}

bool sym27_ready(const lfric_op *op) {
  return op->sym && op->d_sympos && op->to == op->from && role_kind(*op->to) == RoleKind::W0K1;
}

// the order-1 W0 symmetric-packed copy (sym27_apply.h): one unit per (cell, layer), cells in
// cell-id order; the gather columns' cell tables; one right-hand side's row-sum buffer
template <typename T>
static int build_sym27_t(lfric_op *op) {
  const FunctionSpace &f = *op->to;
  const int nc = f.ncells, nl = f.nlayers;
  const long long units = static_cast<long long>(nc) * nl;
  const size_t gss = s27_group_stride<T>();
  SymBuild b;
  b.sym_b = sizeof(T) * gss * static_cast<size_t>((units + 63) / 64);
  b.z_b = sizeof(T) * s27_zalloc(nc, nl, 1);
  b.slotpos_b = sizeof(int) * nc;
  if (int rc = dmalloc(&b.sym, b.sym_b)) return rc;
  if (int rc = dmalloc(&b.z, b.z_b)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&b.flag), sizeof(int))) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&b.slotpos), b.slotpos_b)) return rc;
  std::vector<int> slotpos(nc);
  for (int p = 0; p < nc; ++p) slotpos[f.cell_of_pos[p]] = p;
  LF_HIP(hipMemcpy(b.slotpos, slotpos.data(), b.slotpos_b, hipMemcpyHostToDevice));
  {
    const int ncol = f.gather_ncols;
    std::vector<int> off(ncol + 1), ent;
    LF_HIP(hipMemcpy(off.data(), f.d_goff, sizeof(int) * (ncol + 1), hipMemcpyDeviceToHost));
    ent.resize(off[ncol]);
    LF_HIP(hipMemcpy(ent.data(), f.d_gent, sizeof(int) * ent.size(), hipMemcpyDeviceToHost));
    std::vector<int4> h(ncol);
    for (int c = 0; c < ncol; ++c) {
      LF_CHECK(off[c + 1] - off[c] <= 4, LFRIC_E_UNSUPPORTED,
               "the order-1 symmetric copy needs at most 4 cells around each dof column");
      int u[4] = {-1, -1, -1, -1};
      for (int q = off[c]; q < off[c + 1]; ++q) u[q - off[c]] = (ent[q] >> 5) * 18 + s27_slot(ent[q] & 31);
      h[c] = make_int4(u[0], u[1], u[2], u[3]);
    }
    b.zu_b = sizeof(int4) * h.size();
    if (int rc = dmalloc(reinterpret_cast<void **>(&b.zu), b.zu_b)) return rc;
    LF_HIP(hipMemcpy(b.zu, h.data(), b.zu_b, hipMemcpyHostToDevice));
  }
  int h_flag = 0;
  LF_HIP(hipMemcpyAsync(b.flag, &h_flag, sizeof(int), hipMemcpyHostToDevice, cur_stream()));
  LAUNCH((k_build_sym27<T>), grid_for(units), static_cast<const T *>(op->d), op->gs, f.d_ebase, b.slotpos,
         static_cast<T *>(b.sym), gss, units, nl, op->row_top, b.flag);
  CHECK_LAUNCH();
  LF_HIP(hipMemcpyAsync(&h_flag, b.flag, sizeof(int), hipMemcpyDeviceToHost, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  LF_CHECK(!h_flag, LFRIC_E_UNSUPPORTED, "the operator's element matrices are not symmetric");
  drop_sym(op);
  op->sym = b.sym;
  op->sym_bytes = b.sym_b;
  op->symz = b.z;
  op->symz_bytes = b.z_b;
  op->d_symzu = b.zu;
  op->symzu_bytes = b.zu_b;
  op->d_sympos = b.slotpos;
  op->sympos_bytes = b.slotpos_b;
  op->gss = gss;
  op->sym_seg = 0;
  b.sym = b.z = nullptr;
  b.zu = nullptr;
  b.slotpos = nullptr;
  return LFRIC_OK;
}

int launch_build_sym(lfric_op *op, int levels_per_thread) {
  FunctionSpace &f = *op->to;
  LF_CHECK(op->to == op->from, LFRIC_E_ARG, "the symmetric-packed copy is for square operators");
  if (role_kind(f) == RoleKind::W0K1 && f.ndf == 27) {
    LF_CHECK(levels_per_thread == 0, LFRIC_E_UNSUPPORTED,
             "the order-1 symmetric-packed copy takes one (cell, layer) unit per thread");
    if (int rc = build_gather(f)) return rc;
    return op->precision == LFRIC_R64 ? build_sym27_t<double>(op) : build_sym27_t<float>(op);
  }
  LF_CHECK(role_kind(f) == RoleKind::W0K0 && f.ndf == 8, LFRIC_E_UNSUPPORTED,
           "the symmetric-packed copy is built for W0 operators of order 0 and 1");
  LF_CHECK(levels_per_thread >= 0 && levels_per_thread <= 64, LFRIC_E_ARG, "levels per thread must be 0..64");
  for (int c = 0; c < f.ncells; ++c)
    for (int j = 0; j < 4; ++j)
      LF_CHECK(f.dofmap[static_cast<size_t>(c) * 8 + j + 4] == f.dofmap[static_cast<size_t>(c) * 8 + j] + 1,
               LFRIC_E_UNSUPPORTED, "the symmetric-packed copy needs each top vertex one level above its bottom vertex");
  if (int rc = build_gather(f)) return rc;
  return op->precision == LFRIC_R64 ? build_sym_t<double>(op, levels_per_thread)
                                    : build_sym_t<float>(op, levels_per_thread);
}

int launch_lma_apply_gather_dot(const lfric_op *op, const void *x, void *y, double *partials, const int *skip) {
  return gather_any(op, x, y, GatherExtra{partials, skip});
}

int launch_lma_apply_solver(const lfric_op *op, const void *x, void *y, const int *skip) {
  return ctx().apply_form == 1 ? gather_any(op, x, y, GatherExtra{nullptr, skip}) : launch_lma_apply(op, x, y, skip);
}

int launch_gather_blocks(const lfric_op *op, int *nblocks) {
  if (int rc = build_gather(*op->to)) return rc;
  *nblocks = gather_dot_blocks(op);
  return LFRIC_OK;
}

template <typename T>
static int transpose_t(const lfric_op *op, const T *x, T *y) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  const Roles roles = roles_of(from);
  const size_t S = op->gs;
  const int nl = from.nlayers;
  for (int c = 0; c < from.ncolours; ++c) {
    const int pos0 = from.colour_start[c], npos = from.colour_start[c + 1] - pos0;
    if (npos == 0) continue;
    const int grid = grid_for(static_cast<long long>(npos) * (nl + 1));
#define TR_CASE(A, B)                                                                                        \
  if (to.ndf == A && from.ndf == B) {                                                                       \
    LAUNCH((k_transpose_colour<T, A, B>), grid, static_cast<const T *>(op->d), S, from.d_ebase, x, y, to.d_map,   \
           from.d_map, from.d_first, pos0, npos, nl, roles, op->row_top);                                  \
    continue;                                                                                               \
  }
    TR_CASE(1, 6)
    TR_CASE(8, 8)
    TR_CASE(6, 6)
    TR_CASE(1, 1)
#undef TR_CASE
    return fail(LFRIC_E_UNSUPPORTED, "transpose apply: unsupported (ndf_to, ndf_from)");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_lma_transpose(const lfric_op *op, const void *x, void *y) {
  if (op->precision == LFRIC_R64)
    return transpose_t<double>(op, static_cast<const double *>(x), static_cast<double *>(y));
  if (op->precision == LFRIC_R32)
    return transpose_t<float>(op, static_cast<const float *>(x), static_cast<float *>(y));
  return fail(LFRIC_E_UNSUPPORTED, "transpose apply: precision");
}

template <typename T>
static int dg_t(const lfric_op *op, const T *x, T *y) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  LF_CHECK(to.top_mask == 0, LFRIC_E_ARG, "dg apply needs a discontinuous output space");
  const Roles roles = roles_of(to);
  const size_t S = op->gs;
  const int nl = to.nlayers, npos = to.ncells;
  const int grid = grid_for(static_cast<long long>(npos) * (nl + 1));
#define DG_CASE(A, B)                                                                                            \
  if (to.ndf == A && from.ndf == B) {                                                                           \
    LAUNCH((k_lma_colour<T, A, B, RolesRuntime>), grid, static_cast<const T *>(op->d), S, to.d_ebase, x, y,     \
           to.d_map,                                                                                           \
           from.d_map, static_cast<const uint32_t *>(nullptr), 0, npos, nl, RolesRuntime{roles},               \
           static_cast<const int *>(nullptr), op->xwin);                                                         \
    CHECK_LAUNCH();                                                                                             \
    return LFRIC_OK;                                                                                            \
  }
  DG_CASE(1, 6)
  DG_CASE(1, 1)
  DG_CASE(1, 8)
#undef DG_CASE
  return fail(LFRIC_E_UNSUPPORTED, "dg apply: unsupported (ndf_to, ndf_from)");
}

template <typename T>
static int dg_inc_t(const lfric_op *op, const T *x, T *y) {
  const FunctionSpace &to = *op->to, &from = *op->from;
  LF_CHECK(to.top_mask == 0, LFRIC_E_ARG, "dg_inc apply needs a discontinuous output space");
  LF_CHECK(to.nlayers == from.nlayers && to.ncells == from.ncells, LFRIC_E_ARG, "spaces of different meshes");
  const int npos = to.ncells, nl = to.nlayers;
  const int grid = grid_for(static_cast<long long>(npos) * nl);
#define DGI_CASE(A, B)                                                                                             \
  if (to.ndf == A && from.ndf == B) {                                                                             \
    LAUNCH((k_dg_inc<T, A, B>), grid, static_cast<const T *>(op->d), op->gs, to.d_ebase, x, y, to.d_map, from.d_map, \
           npos, nl, A, B);                                                                                       \
    CHECK_LAUNCH();                                                                                               \
    return LFRIC_OK;                                                                                              \
  }
  DGI_CASE(1, 6)
  DGI_CASE(1, 1)
  DGI_CASE(1, 8)
#undef DGI_CASE
  LAUNCH((k_dg_inc<T, 0, 0>), grid, static_cast<const T *>(op->d), op->gs, to.d_ebase, x, y, to.d_map, from.d_map, npos,
         nl, to.ndf, from.ndf);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_dg_inc_apply(const lfric_op *op, const void *x, void *y) {
  if (op->precision == LFRIC_R64) return dg_inc_t<double>(op, static_cast<const double *>(x), static_cast<double *>(y));
  if (op->precision == LFRIC_R32) return dg_inc_t<float>(op, static_cast<const float *>(x), static_cast<float *>(y));
  return fail(LFRIC_E_UNSUPPORTED, "dg_inc apply: precision");
}

int launch_enforce_bc(const FunctionSpace &f, int prec, void *y, const int *skip) {
  if (!f.bc_bottom && !f.bc_top) return LFRIC_OK;
  const int grid = grid_for(f.ncells);
  if (prec == LFRIC_R64)
    LAUNCH(k_enforce_bc<double>, grid, static_cast<double *>(y), f.d_map, f.ncells, f.ndf, f.nlayers, f.bc_bottom,
           f.bc_top, skip);
  else if (prec == LFRIC_R32)
    LAUNCH(k_enforce_bc<float>, grid, static_cast<float *>(y), f.d_map, f.ncells, f.ndf, f.nlayers, f.bc_bottom,
           f.bc_top, skip);
  else
    return fail(LFRIC_E_UNSUPPORTED, "enforce_bc: precision");
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_dg_apply(const lfric_op *op, const void *x, void *y) {
  if (op->precision == LFRIC_R64) return dg_t<double>(op, static_cast<const double *>(x), static_cast<double *>(y));
  if (op->precision == LFRIC_R32) return dg_t<float>(op, static_cast<const float *>(x), static_cast<float *>(y));
  return fail(LFRIC_E_UNSUPPORTED, "dg apply: precision");
}

template <typename T>
static int diag_t(const lfric_op *op, T *d) {
  const FunctionSpace &f = *op->to;
  LF_CHECK(op->to == op->from, LFRIC_E_ARG, "mm_diagonal needs a square operator on one space");
  const Roles roles = roles_of(f);
  const size_t S = op->gs;
  for (int c = 0; c < f.ncolours; ++c) {
    const int pos0 = f.colour_start[c], npos = f.colour_start[c + 1] - pos0;
    if (npos == 0) continue;
    const int grid = grid_for(static_cast<long long>(npos) * (f.nlayers + 1));
#define DI_CASE(A)                                                                                                    \
  if (f.ndf == A) {                                                                                                  \
    LAUNCH((k_diag_colour<T, A>), grid, static_cast<const T *>(op->d), S, f.d_ebase, d, f.d_map, f.d_first, pos0,     \
           npos, f.nlayers,                                                                                          \
           roles);                                                                                                   \
    continue;                                                                                                        \
  }
    DI_CASE(8)
    DI_CASE(27)
    DI_CASE(6)
    DI_CASE(1)
#undef DI_CASE
    return fail(LFRIC_E_UNSUPPORTED, "mm_diagonal: unsupported ndf");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_mm_diagonal(const lfric_op *op, void *diag) {
  if (op->precision == LFRIC_R64) return diag_t<double>(op, static_cast<double *>(diag));
  if (op->precision == LFRIC_R32) return diag_t<float>(op, static_cast<float *>(diag));
  return fail(LFRIC_E_UNSUPPORTED, "mm_diagonal: precision");
}

// ---- vectors -------------------------------------------------------------------------------
static inline int vgrid(long long n) { return grid_for(n, kBlock, 8192); }

int launch_set(int prec, void *v, long long n, double s) {
  if (prec == LFRIC_R64) {
    LAUNCH(k_set<double>, vgrid(n), static_cast<double *>(v), n, s);
  } else if (prec == LFRIC_R32) {
    LAUNCH(k_set<float>, vgrid(n), static_cast<float *>(v), n, static_cast<float>(s));
  } else {
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_axpy(int prec, void *y, const void *x, long long n, double a) {
  if (prec == LFRIC_R64) {
    LAUNCH(k_axpy<double>, vgrid(n), static_cast<double *>(y), static_cast<const double *>(x), n, a);
  } else if (prec == LFRIC_R32) {
    LAUNCH(k_axpy<float>, vgrid(n), static_cast<float *>(y), static_cast<const float *>(x), n, static_cast<float>(a));
  } else {
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_aypx(int prec, void *y, const void *x, long long n, double a) {
  if (prec == LFRIC_R64) {
    LAUNCH(k_aypx<double>, vgrid(n), static_cast<double *>(y), static_cast<const double *>(x), n, a);
  } else if (prec == LFRIC_R32) {
    LAUNCH(k_aypx<float>, vgrid(n), static_cast<float *>(y), static_cast<const float *>(x), n, static_cast<float>(a));
  } else {
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_axpby(int prec, void *z, double a, const void *x, double b, const void *y, long long n) {
  if (prec == LFRIC_R64) {
    LAUNCH(k_axpby<double>, vgrid(n), static_cast<double *>(z), a, static_cast<const double *>(x), b,
           static_cast<const double *>(y), n);
  } else if (prec == LFRIC_R32) {
    LAUNCH(k_axpby<float>, vgrid(n), static_cast<float *>(z), static_cast<float>(a), static_cast<const float *>(x),
           static_cast<float>(b), static_cast<const float *>(y), n);
  } else {
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_scale(int prec, void *v, long long n, double a) {
  if (prec == LFRIC_R64) {
    LAUNCH(k_scale<double>, vgrid(n), static_cast<double *>(v), n, a);
  } else if (prec == LFRIC_R32) {
    LAUNCH(k_scale<float>, vgrid(n), static_cast<float *>(v), n, static_cast<float>(a));
  } else {
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_copy(int prec, void *dst, const void *src, long long n) {
  const size_t es = prec == LFRIC_R64 ? 8 : 4;
  LF_HIP(hipMemcpyAsync(dst, src, es * static_cast<size_t>(n), hipMemcpyDeviceToDevice, cur_stream()));
  return LFRIC_OK;
}

int launch_divide(int prec, void *z, const void *x, const void *d, long long n) {
  if (prec == LFRIC_R64)
    LAUNCH(k_divide<double>, vgrid(n), static_cast<double *>(z), static_cast<const double *>(x),
           static_cast<const double *>(d), n);
  else if (prec == LFRIC_R32)
    LAUNCH(k_divide<float>, vgrid(n), static_cast<float *>(z), static_cast<const float *>(x),
           static_cast<const float *>(d), n);
  else
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_axpy_mixed(double *y, double a, int prec_x, const void *x, long long n) {
  if (prec_x == LFRIC_R64)
    LAUNCH(k_axpy_mixed<double>, vgrid(n), y, a, static_cast<const double *>(x), n);
  else if (prec_x == LFRIC_R32)
    LAUNCH(k_axpy_mixed<float>, vgrid(n), y, a, static_cast<const float *>(x), n);
  else
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_dot(int prec, const void *x, const void *y, long long n, double *d_out) {
  Context &c = ctx();
  if (prec == LFRIC_R64)
    LAUNCH(k_dot<double>, kRedBlocks, static_cast<const double *>(x), static_cast<const double *>(y), n, c.d_partials,
           c.d_counter, d_out);
  else if (prec == LFRIC_R32)
    LAUNCH(k_dot<float>, kRedBlocks, static_cast<const float *>(x), static_cast<const float *>(y), n, c.d_partials,
           c.d_counter, d_out);
  else
    return fail(LFRIC_E_UNSUPPORTED, "vector precision");
  LAUNCH((k_finish<1, kEpiDot>), 1, c.d_partials, kRedBlocks, static_cast<SolverState *>(nullptr), d_out, 0);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

// ---- CG ---------------------------------------------------------------------------------------
int launch_cg_residual(int prec, void *r, const void *b, long long n) {
  if (prec == LFRIC_R64)
    LAUNCH(k_cg_residual<double>, vgrid(n), static_cast<double *>(r), static_cast<const double *>(b), n);
  else
    LAUNCH(k_cg_residual<float>, vgrid(n), static_cast<float *>(r), static_cast<const float *>(b), n);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_cg_init_norm(int prec, const void *r, long long n, SolverState *st) {
  if (prec == LFRIC_R64)
    LAUNCH(k_cg_init_norm<double>, kRedBlocks, static_cast<const double *>(r), n, st, ctx().d_partials);
  else
    LAUNCH(k_cg_init_norm<float>, kRedBlocks, static_cast<const float *>(r), n, st, ctx().d_partials);
  LAUNCH((k_finish<1, kEpiInitNorm>), 1, ctx().d_partials, kRedBlocks, st, static_cast<double *>(nullptr), 0);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_cg_init_prec(int prec, const void *r, void *z, void *p, const void *diag, long long n, SolverState *st) {
  if (prec == LFRIC_R64)
    LAUNCH(k_cg_init_prec<double>, kRedBlocks, static_cast<const double *>(r), static_cast<double *>(z),
           static_cast<double *>(p), static_cast<const double *>(diag), n, st, ctx().d_partials);
  else
    LAUNCH(k_cg_init_prec<float>, kRedBlocks, static_cast<const float *>(r), static_cast<float *>(z),
           static_cast<float *>(p), static_cast<const float *>(diag), n, st, ctx().d_partials);
  LAUNCH((k_finish<1, kEpiInitPrec>), 1, ctx().d_partials, kRedBlocks, st, static_cast<double *>(nullptr), 0);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

// partial sums of p.(A p), one per block, into partials (kRedBlocks of them)
int launch_cg_dot_pz(int prec, const void *p, const void *z, long long n, SolverState *st, double *partials) {
  if (prec == LFRIC_R64)
    LAUNCH(k_cg_dot_pz<double>, kRedBlocks, static_cast<const double *>(p), static_cast<const double *>(z), n, st,
           partials);
  else
    LAUNCH(k_cg_dot_pz<float>, kRedBlocks, static_cast<const float *>(p), static_cast<const float *>(z), n, st,
           partials);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

// alpha from partial sums of p.(A p) (one-block finish, for applies with more partials than
// the folded update adds up itself)
int launch_cg_finish_pz(const double *partials, int nblocks, SolverState *st, int par) {
  LAUNCH((k_finish<1, kEpiDotPz>), 1, partials, nblocks, st, static_cast<double *>(nullptr), par);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

// the update's (r.r, r.z) partials: past the first kRedBlocks partials, which hold p.(A p)
static double *update_partials() { return ctx().d_partials + 2 * kRedBlocks; }

int launch_cg_update(int prec, void *x, void *r, void *z, const void *p, const void *diag, long long n,
                     SolverState *st, const double *pz_partials, int npz, int par) {
  double *out = update_partials();
  if (pz_partials) {
    if (prec == LFRIC_R64)
      LAUNCH(k_cg_update_fold<double>, kRedBlocks, static_cast<double *>(x), static_cast<double *>(r),
             static_cast<double *>(z), static_cast<const double *>(p), static_cast<const double *>(diag), n, st,
             pz_partials, npz, out, par);
    else
      LAUNCH(k_cg_update_fold<float>, kRedBlocks, static_cast<float *>(x), static_cast<float *>(r),
             static_cast<float *>(z), static_cast<const float *>(p), static_cast<const float *>(diag), n, st,
             pz_partials, npz, out, par);
  } else if (prec == LFRIC_R64) {
    LAUNCH(k_cg_update<double>, kRedBlocks, static_cast<double *>(x), static_cast<double *>(r), static_cast<double *>(z),
           static_cast<const double *>(p), static_cast<const double *>(diag), n, st, out);
  } else {
    LAUNCH(k_cg_update<float>, kRedBlocks, static_cast<float *>(x), static_cast<float *>(r), static_cast<float *>(z),
           static_cast<const float *>(p), static_cast<const float *>(diag), n, st, out);
  }
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_cg_aypx(int prec, void *p, const void *z, long long n, SolverState *st, int par) {
  // kRedBlocks blocks (grid stride): each block adds the update's partials once
  if (prec == LFRIC_R64)
    LAUNCH(k_cg_aypx_fold<double>, kRedBlocks, static_cast<double *>(p), static_cast<const double *>(z), n, st,
           update_partials(), kRedBlocks, par);
  else
    LAUNCH(k_cg_aypx_fold<float>, kRedBlocks, static_cast<float *>(p), static_cast<const float *>(z), n, st,
           update_partials(), kRedBlocks, par);
  CHECK_LAUNCH();
  return LFRIC_OK;
}

bool helmholtz_fusable(const lfric_linop *L) {
  const FunctionSpace &w3 = *L->D->to, &w2 = *L->D->from;
  return L->kind == 1 && L->M3->to == L->D->to && L->M3->from == L->D->to && w3.ndf == 1 && w2.ndf == 6 &&
         w3.top_mask == 0 && role_kind(w2) == RoleKind::W2K0 && w2.ncells == w3.ncells && w2.nlayers == w3.nlayers &&
         w2.ebase == w3.ebase && (L->precision == LFRIC_R64 || L->precision == LFRIC_R32);
}

// the composite's column table for k_build_dtc and k_helm_transpose_div_dtc (the gather lists
// of the W2 space resolved to ebase and W3 dof per cell); none when a column has more than two
// cells
static int build_helm_table(const lfric_linop *L) {
  if (L->hcol_state) return LFRIC_OK;
  const FunctionSpace &w2 = *L->D->from, &w3 = *L->D->to;
  struct Item {
    int base, pos, df;
  };
  std::vector<Item> items;
  for (int p = 0; p < w2.ncells; ++p)
    for (int d = 0; d < 6; ++d)
      if (!((w2.top_mask >> d) & 1u)) items.push_back({w2.dofmap[static_cast<size_t>(w2.cell_of_pos[p]) * 6 + d] - 1, p, d});
  std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.base < b.base; });
  std::vector<int4> tab;
  bool ok = true;
  for (size_t q = 0; q < items.size() && ok; ++q) {
    const Item &it = items[q];
    const int ent = w2.ebase[it.pos], bj = (w3.dofmap[w3.cell_of_pos[it.pos]] - 1) * 8 + it.df;
    if (q == 0 || it.base != items[q - 1].base) {
      tab.push_back(make_int4(ent, bj, -1, 0));
    } else if (tab.back().z < 0) {
      tab.back().z = ent;
      tab.back().w = bj;
    } else {
      ok = false;
    }
  }
  ok = ok && static_cast<int>(tab.size()) == w2.gather_ncols && static_cast<long long>(w3.undf) * 8 < (1LL << 31) &&
       static_cast<long long>(w2.gather_ncols) * (w2.nlayers + 1) + kBlock < (1LL << 31);
  L->hcol_state = -1;
  if (!ok) return LFRIC_OK;
  L->hcol_bytes = sizeof(int4) * tab.size();
  if (int rc = dmalloc(reinterpret_cast<void **>(&L->d_hcol), L->hcol_bytes)) return rc;
  LF_HIP(hipMemcpy(L->d_hcol, tab.data(), L->hcol_bytes, hipMemcpyHostToDevice));
  L->hcol_state = 1;
  return LFRIC_OK;
}

template <typename T>
static int helmholtz_t(const lfric_linop *L, const T *x, T *y) {
  FunctionSpace &w2 = *L->D->from;
  const FunctionSpace &w3 = *L->D->to;
  if (int rc = build_gather(w2)) return rc;
  if (int rc = build_helm_table(L)) return rc;
  T *t = static_cast<T *>(L->t_w2);
  const int tgrid = grid_for(static_cast<long long>(w2.gather_ncols) * (w2.nlayers + 1));
  if (L->hcol_state == 1 && L->dtc_version != L->D->version) {  // (re)build the copy of D
    if (!L->d_dtc) {
      L->dtc_bytes = 2 * sizeof(T) * static_cast<size_t>(w2.undf);
      if (int rc = dmalloc(&L->d_dtc, L->dtc_bytes)) return rc;
    }
    LAUNCH((k_build_dtc<T>), tgrid, static_cast<const T *>(L->D->d), L->D->gs, w2.d_gcol,
           static_cast<const int4 *>(L->d_hcol), w2.gather_ncols, w2.nlayers, static_cast<T *>(L->d_dtc));
    CHECK_LAUNCH();
    L->dtc_version = L->D->version;
  }
  if (L->hcol_state == 1)
    LAUNCH((k_helm_transpose_div_dtc<T>), tgrid, static_cast<const T *>(L->d_dtc), x,
           static_cast<const T *>(L->m2->d), t, w2.d_gcol, static_cast<const int4 *>(L->d_hcol), w2.gather_ncols,
           w2.nlayers);
  else
    LAUNCH((k_helm_transpose_div<T>), tgrid, static_cast<const T *>(L->D->d), L->D->gs, w3.d_ebase, x, w3.d_map,
           static_cast<const T *>(L->m2->d), t, w2.d_gcol, w2.d_goff, w2.d_gent, w2.gather_ncols, w2.nlayers,
           roles_of(w2));
  const int ugrid = grid_for(static_cast<long long>(w3.ncells) * w3.nlayers);
  if (L->hcol_state == 1 && (L->drc_dver != L->D->version || L->drc_mver != L->M3->version)) {
    if (!L->d_drc) {  // (re)build the compact rows of D and M3
      L->drc_bytes = 7 * sizeof(T) * static_cast<size_t>(w3.ncells) * w3.nlayers;
      if (int rc = dmalloc(&L->d_drc, L->drc_bytes)) return rc;
    }
    LAUNCH((k_build_drc<T>), ugrid, static_cast<const T *>(L->D->d), L->D->gs, static_cast<const T *>(L->M3->d),
           L->M3->gs, w3.d_ebase, w3.ncells, w3.nlayers, static_cast<T *>(L->d_drc));
    CHECK_LAUNCH();
    L->drc_dver = L->D->version;
    L->drc_mver = L->M3->version;
  }
  if (L->hcol_state == 1)
    LAUNCH((k_helm_dg_m3_drc<T>), ugrid, static_cast<const T *>(L->d_drc), t, w2.d_map, x, w3.d_map, y, w3.ncells,
           w3.nlayers, static_cast<T>(L->tau));
  else
    LAUNCH((k_helm_dg_m3<T>), ugrid, static_cast<const T *>(L->D->d), L->D->gs, static_cast<const T *>(L->M3->d),
           L->M3->gs, w3.d_ebase, t, w2.d_map, x, w3.d_map, y, w3.ncells, w3.nlayers, static_cast<T>(L->tau));
  CHECK_LAUNCH();
  return LFRIC_OK;
}

int launch_helmholtz_fused(const lfric_linop *L, const void *x, void *y) {
  if (L->precision == LFRIC_R64)
    return helmholtz_t<double>(L, static_cast<const double *>(x), static_cast<double *>(y));
  return helmholtz_t<float>(L, static_cast<const float *>(x), static_cast<float *>(y));
}

}  // namespace lfric
