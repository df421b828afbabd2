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

}  // namespace
}  // namespace lfric
