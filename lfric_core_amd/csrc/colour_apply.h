// Column roles of the element types with compile-time structure, and the colour-ordered
// multi-RHS unit shared by the single-operator kernels and the batch.
//
// Level-aligned units: unit (cell position, level m) owns, for every non-top dof i of the cell,
// the dof at level m of i's column.  Its value is row i of the layer-m element (when m < nl)
// plus, for a dof with a top partner pt, row pt of the layer-(m-1) element (when m > 0); both
// rows are stored at level-aligned slot e = ebase[pos] + m.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"

namespace lfric {
namespace {

struct RolesNone {
  static constexpr bool kStatic = true;
  __host__ __device__ static constexpr bool top(int) { return false; }
  __host__ __device__ static constexpr int partner(int) { return -1; }
};
struct RolesW0K0 {  // 4 bottom vertices, 4 top vertices
  static constexpr bool kStatic = true;
  __host__ __device__ static constexpr bool top(int i) { return i >= 4; }
  __host__ __device__ static constexpr int partner(int i) { return i < 4 ? i + 4 : -1; }
};
struct RolesW2K0 {  // faces W,S,E,N (height nl), B -> T
  static constexpr bool kStatic = true;
  __host__ __device__ static constexpr bool top(int i) { return i == 5; }
  __host__ __device__ static constexpr int partner(int i) { return i == 4 ? 5 : -1; }
};
struct RolesW0K1 {  // B->T face, WB..NB -> WT..NT edges, bottom -> top vertices
  static constexpr bool kStatic = true;
  __host__ __device__ static constexpr bool top(int i) { return i == 6 || (i >= 15 && i <= 18) || i >= 23; }
  __host__ __device__ static constexpr int partner(int i) {
    return i == 5 ? 6 : (i >= 7 && i <= 10) ? i + 8 : (i >= 19 && i <= 22) ? i + 4 : -1;
  }
};
template <typename T, int NR>
struct RVec {
  typedef T type __attribute__((ext_vector_type(NR)));
};

// multi-RHS vectors hold NR values per dof, RHS fastest: (dof, r) at dof*NR + r
template <typename T, int N2, int NR>
__device__ inline void row_dot_mrhs(const T *__restrict__ A, size_t gs, long long e, int r, const T *__restrict__ x,
                                    const int (&b)[N2], int lev, typename RVec<T, NR>::type &acc) {
  typedef typename RVec<T, NR>::type VR;
  for_row<T, N2>(A, gs, e, r, [&](int j, T a) {
    const VR xv = *reinterpret_cast<const VR *>(x + (static_cast<size_t>(b[j]) + lev) * NR);
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = fma(a, xv[k], acc[k]);
  });
}

// One unit, NR right-hand sides: c = row i . x(level m), cp = row pt . x(level m-1), each an
// fma chain over the row in df2 order from zero; y (=|+=) c + cp (= where the unit's colour
// touches the dof first).  When the gathered x of one level fits in 128 registers and the roles
// are static, each level's x is loaded once for all rows (instead of once per row); the sums are
// the same in the same order either way.
template <typename T, int N1, int N2, int NR, class R>
__device__ inline void colour_unit_mrhs(const T *__restrict__ A, size_t gs, long long e, const int *__restrict__ map1p,
                                        const int *__restrict__ map2p, uint32_t ft, const T *__restrict__ x,
                                        T *__restrict__ y, int m, int nl, R roles) {
  typedef typename RVec<T, NR>::type VR;
  const bool has_m = m < nl, has_p = m > 0;
  int b[N2];
#pragma unroll
  for (int j = 0; j < N2; ++j) b[j] = map2p[j];
  if constexpr (R::kStatic && sizeof(T) * N2 * NR <= 512) {
    VR X[N2];
    VR cp[N1];
    if (has_p) {
#pragma unroll
      for (int j = 0; j < N2; ++j) X[j] = *reinterpret_cast<const VR *>(x + (static_cast<size_t>(b[j]) + m - 1) * NR);
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        if (roles.top(i) || roles.partner(i) < 0) continue;
        cp[i] = VR{};
        for_row<T, N2>(A, gs, e, roles.partner(i), [&](int j, T a) {
#pragma unroll
          for (int k = 0; k < NR; ++k) cp[i][k] = fma(a, X[j][k], cp[i][k]);
        });
      }
    }
    if (has_m) {
#pragma unroll
      for (int j = 0; j < N2; ++j) X[j] = *reinterpret_cast<const VR *>(x + (static_cast<size_t>(b[j]) + m) * NR);
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      if (roles.top(i)) continue;
      const int pt = roles.partner(i);
      if (!has_m && pt < 0) continue;
      VR c = {};
      if (has_m)
        for_row<T, N2>(A, gs, e, i, [&](int j, T a) {
#pragma unroll
          for (int k = 0; k < NR; ++k) c[k] = fma(a, X[j][k], c[k]);
        });
      VR out;
      if (pt >= 0) {
        const VR p = has_p ? cp[i] : VR{};
#pragma unroll
        for (int k = 0; k < NR; ++k) out[k] = c[k] + p[k];
      } else {
        out = c;
      }
      VR *yd = reinterpret_cast<VR *>(y + (static_cast<size_t>(map1p[i]) + m) * NR);
      if ((ft >> i) & 1u)
        *yd = out;
      else
        *yd = *yd + out;
    }
  } else {
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      if (roles.top(i)) continue;
      const int pt = roles.partner(i);
      if (!has_m && pt < 0) continue;
      VR c = {}, cp = {};
      if (has_m) row_dot_mrhs<T, N2, NR>(A, gs, e, i, x, b, m, c);
      if (pt >= 0 && has_p) row_dot_mrhs<T, N2, NR>(A, gs, e, pt, x, b, m - 1, cp);
      VR out;
#pragma unroll
      for (int k = 0; k < NR; ++k) out[k] = pt >= 0 ? c[k] + cp[k] : c[k];
      VR *yd = reinterpret_cast<VR *>(y + (static_cast<size_t>(map1p[i]) + m) * NR);
      if ((ft >> i) & 1u)
        *yd = out;
      else
        *yd = *yd + out;
    }
  }
}

// The row-gathered copy of a square operator whose dof columns each take rows from at most
// kRowsKmax cells: the rows that gather point t = (column, level m) reads are stored together,
// slot s = 2k + h holding, for the k-th cell of the column's list, its top-partner row of layer
// m-1 (h = 0) or its bottom row of layer m (h = 1), each padded to RS entries (whole chunks).
// Entry j of slot s of point t: (t/64)*gsr + ((s*RS + j)/V*64 + t%64)*V + (s*RS + j)%V.
constexpr int kRowsKmax = 4;
template <typename T, int N2>
constexpr int rows_rs() {
  return (N2 + chunk_entries<T>() - 1) / chunk_entries<T>() * chunk_entries<T>();
}

// Gather point t = (column, level m) of a lowest-order W0 operator over its row-gathered copy,
// NR right-hand sides: per cell of the column's list the top-partner row of layer m-1, then the
// bottom row of layer m, all in one fma chain per RHS (the gather form's order).  The cell's
// top vertex j+4 lies one level above its bottom vertex j, so the x entries are the cell's 4
// columns at levels m-1, m and m+1, each loaded once.  Nothing past the end.
template <typename T, int NR>
__device__ inline void rows_point_w0_mrhs(const T *__restrict__ rows, size_t gsr, const T *__restrict__ x,
                                          T *__restrict__ y, const int *__restrict__ map2,
                                          const int2 *__restrict__ cols, const int *__restrict__ off,
                                          const int *__restrict__ ent, int ncols, int nl, long long t) {
  typedef typename RVec<T, NR>::type VR;
  constexpr int V = chunk_entries<T>(), RS = rows_rs<T, 8>();
  typedef T VecT __attribute__((ext_vector_type(V)));
  const int H = nl + 1;
  if (t >= static_cast<long long>(ncols) * H) return;
  const int col = static_cast<int>(t / H);
  const int m = static_cast<int>(t - static_cast<long long>(col) * H);
  const int2 ci = cols[col];
  if (m == nl && !ci.y) return;
  const bool has_m = m < nl, has_p = m > 0;
  const int q0 = off[col], cnt = off[col + 1] - q0;
  const T *Rt = rows + static_cast<size_t>(t >> 6) * gsr + static_cast<size_t>(t & 63) * V;
  VR acc = {};
#pragma unroll
  for (int k = 0; k < kRowsKmax; ++k) {
    if (k >= cnt) break;
    const int pos = ent[q0 + k] >> 5;
    int b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = map2[static_cast<size_t>(pos) * 8 + j];
    VR xl[3][4];  // x of the cell's 4 columns at levels m-1, m, m+1
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      xl[0][j] = has_p ? *reinterpret_cast<const VR *>(x + (static_cast<size_t>(b[j]) + m - 1) * NR) : VR{};
      xl[1][j] = *reinterpret_cast<const VR *>(x + (static_cast<size_t>(b[j]) + m) * NR);
      xl[2][j] = has_m ? *reinterpret_cast<const VR *>(x + (static_cast<size_t>(b[j]) + m + 1) * NR) : VR{};
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
          const VR &xv = j < 4 ? xl[h][j] : xl[h + 1][j - 4];
#pragma unroll
          for (int r = 0; r < NR; ++r) acc[r] = fma(v[q], xv[r], acc[r]);
        }
      }
    }
  }
  *reinterpret_cast<VR *>(y + (static_cast<size_t>(ci.x) + m) * NR) = acc;
}

}  // namespace
}  // namespace lfric
