// Symmetric-packed apply of square order-1 W0 operators (27 x 27 element matrices, Q2 nodes),
// one or several right-hand sides, in two phases; device code shared by the single-operator
// kernels and the batch.
//
// The packed copy holds, per (cell, layer) unit, the upper triangle of the unit's element: 378
// entries row by row, (0,0) (0,1) .. (0,26) (1,1) .. (26,26).  Unit u = k*nl + m is layer m of
// the k-th cell in cell-id order (neighbouring units share x entries in one L2); entry q of unit
// u sits at (u/64)*gss + ((q/V)*64 + u%64)*V + q%V, so that a lane's V consecutive entries are
// one 16-byte load and the lanes of consecutive units read one contiguous block.
//
// Phase 1 (sym27_unit): a unit streams its triangle once.  Entry (i, j) adds a_ij x_j to row i's
// chain and, for j > i, a_ij x_i to row j's chain, so that every row's chain takes its terms in
// column order 0..26 from zero: the fma chain of that row of the full element (a_ji = a_ij bit
// for bit, which building the copy checks).  The 9 top rows' chains belong to the dofs of the
// unit one layer up, which the next lanes hold: they pass up by one shuffle.  A wavefront takes
// 64/L - 1 new units and repeats, in its first L lanes, the previous wavefront's last unit (whose
// outputs it does not write), so that every unit has its lower neighbour in its wavefront.  A
// unit writes, for each of its 18 non-top dofs i (slot r = s27_slot(i)), the colour-ordered
// unit's output: row i's chain, plus for a bottom dof the top partner row of the layer below
// (+ 0 in the bottom layer), at z[((pos*18 + r)*(nl+1) + m)*NR + rhs]; the top layer's unit
// also writes 0 + its top rows' chains at level nl.
//
// Phase 2 (sym27_gather_point): the dof at level m of a dof column is the sum of its cells'
// outputs in colour order, the first written and the others added: the colour-ordered form's
// arithmetic (colour_unit_mrhs), so the result equals the colour form's bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "colour_apply.h"
#include "device_common.h"

namespace lfric {
namespace {

constexpr int kS27 = 27;
constexpr int kS27Tri = kS27 * (kS27 + 1) / 2;  // 378

__host__ __device__ constexpr int s27_index(int i, int j) { return i * kS27 - i * (i - 1) / 2 + (j - i); }
__host__ __device__ constexpr int s27_row(int q) {
  int i = 0;
  while (q >= kS27 - i) {
    q -= kS27 - i;
    ++i;
  }
  return i;
}
__host__ __device__ constexpr int s27_col(int q) { return q - s27_index(s27_row(q), s27_row(q)) + s27_row(q); }

template <typename T>
constexpr int s27_chunks() {
  return (kS27Tri + chunk_entries<T>() - 1) / chunk_entries<T>();
}
// values per 64-unit group of the packed copy
template <typename T>
constexpr size_t s27_group_stride() {
  return static_cast<size_t>(64) * s27_chunks<T>() * chunk_entries<T>();
}
// right-hand sides per lane in phase 1: a unit's NR right-hand sides are spread over NR/NRL
// consecutive lanes, which read the same triangle entries (FP32: two per lane keeps the 27 x
// values and 27 chains of a lane within about 170 registers)
template <typename T, int NR>
constexpr int s27_nrl() {
  return sizeof(T) == 8 ? 1 : (NR < 2 ? NR : 2);
}

// 16-byte chunks of the triangle in flight per lane in phase 1 (measured on C48x70 and C96x70
// and the batched configuration: 8 with one right-hand side per lane, 12 with two)
template <int NRL>
constexpr int s27_prefetch() {
  return NRL == 1 ? 8 : 12;
}

template <int E, int N, class F>
__device__ __forceinline__ void s27_for(F &&f) {
  if constexpr (E < N) {
    f(std::integral_constant<int, E>{});
    s27_for<E + 1, N>(f);
  }
}

// slot of non-top dof i among the 18 non-top dofs of an order-1 W0 cell (-1 for a top dof)
__host__ __device__ constexpr int s27_slot(int i) {
  return RolesW0K1::top(i) ? -1 : i - (i > 6) - (i > 18 ? 4 : 0);
}

// bottom partner of top dof j of an order-1 W0 cell
__host__ __device__ constexpr int s27_bottom_of(int j) { return j == 6 ? 5 : j <= 18 ? j - 8 : j - 4; }
static_assert(RolesW0K1::partner(s27_bottom_of(6)) == 6 && RolesW0K1::partner(s27_bottom_of(15)) == 15 &&
                  RolesW0K1::partner(s27_bottom_of(18)) == 18 && RolesW0K1::partner(s27_bottom_of(23)) == 23 &&
                  RolesW0K1::partner(s27_bottom_of(26)) == 26,
              "bottom partners of the top dofs");
static_assert(s27_slot(0) == 0 && s27_slot(5) == 5 && s27_slot(7) == 6 && s27_slot(14) == 13 && s27_slot(19) == 14 &&
                  s27_slot(22) == 17,
              "slots of the non-top dofs");

// values of z for ncells cells of nl layers and NR right-hand sides (18 slots of nl+1 levels per
// cell), and the junk rows after them that the lanes of unwritten units store to
__host__ __device__ inline size_t s27_zvalues(long long ncells, int nl, int NR) {
  return static_cast<size_t>(ncells) * 18 * (nl + 1) * NR;
}
__host__ __device__ inline size_t s27_zalloc(long long ncells, int nl, int NR) {
  return s27_zvalues(ncells, nl, NR) + static_cast<size_t>(18) * (nl + 1) * NR;
}

// threads of phase 1 for `units` units: whole wavefronts of 64/L - 1 new units each
template <typename T, int NR>
constexpr long long s27_threads(long long units) {
  constexpr int L = NR / s27_nrl<T, NR>(), W = 64 / L - 1;
  return (units + W - 1) / W * 64;
}

// Phase 1 for thread g of a problem (g < s27_threads(units); every lane of a wavefront must get
// here, for the shuffle).  units = ncells*nl; slotpos[k]: the k-th cell's colour-major position;
// map2: level-aligned dof bases per position.
template <typename T, int NR, int NRL>
__device__ inline void sym27_unit(const T *__restrict__ sym, size_t gss, const T *__restrict__ x,
                                  const int *__restrict__ map2, const int *__restrict__ slotpos, T *__restrict__ z,
                                  long long units, int nl, long long g) {
  constexpr int L = NR / NRL, V = chunk_entries<T>(), NCH = s27_chunks<T>(), D = s27_prefetch<NRL>();
  constexpr int W = 64 / L - 1;
  static_assert(NR % NRL == 0 && 64 % L == 0, "right-hand sides per lane must divide the unit's");
  typedef T VecT __attribute__((ext_vector_type(V)));
  typedef typename RVec<T, NRL>::type VR;
  const int lane = static_cast<int>(g & 63), slot = lane / L, rl = lane - slot * L;
  const long long u = (g >> 6) * W + slot - 1;
  const bool valid = u >= 0 && u < units;
  const long long uc = valid ? u : 0;  // loads of a lane past the ends read unit 0
  const int k = static_cast<int>(uc / nl), m = static_cast<int>(uc - static_cast<long long>(k) * nl);
  const T *St = sym + static_cast<size_t>(uc >> 6) * gss + static_cast<size_t>(uc & 63) * V;
  VecT buf[D];
#pragma unroll
  for (int c = 0; c < D; ++c)
    buf[c] = c < NCH ? __builtin_nontemporal_load(reinterpret_cast<const VecT *>(St + static_cast<size_t>(c) * 64 * V))
                     : VecT{};
  const int pos = slotpos[k];
  const int *mp = map2 + static_cast<size_t>(pos) * kS27;
  VR X[kS27], c[kS27];
#pragma unroll
  for (int j = 0; j < kS27; ++j) {
    X[j] = *reinterpret_cast<const VR *>(x + (static_cast<size_t>(mp[j]) + m) * NR + rl * NRL);
    c[j] = VR{};
  }
  const int H = nl + 1;
  // a lane whose unit is not written (a repeated unit, or past the end) writes to the junk rows
  // past the end of z instead: a store under a per-lane branch ahead of the shuffles below makes
  // the compiler keep the whole triangle in registers
  const bool store = valid && slot > 0;
  T *zb = z + (store ? (static_cast<size_t>(pos) * 18 * H + m) * NR : s27_zvalues(units / nl, nl, NR)) + rl * NRL;
  VecT cur{};
  s27_for<0, kS27Tri>([&](auto ec) {
    constexpr int e = decltype(ec)::value;
    constexpr int i = s27_row(e), j = s27_col(e), ch = e / V, q = e % V;
    if constexpr (q == 0) {
      cur = buf[ch % D];
      if constexpr (ch + D < NCH)
        buf[ch % D] =
            __builtin_nontemporal_load(reinterpret_cast<const VecT *>(St + static_cast<size_t>(ch + D) * 64 * V));
    }
    const T a = cur[q];
#pragma unroll
    for (int r = 0; r < NRL; ++r) c[i][r] = fma(a, X[j][r], c[i][r]);
    if constexpr (j > i) {
#pragma unroll
      for (int r = 0; r < NRL; ++r) c[j][r] = fma(a, X[i][r], c[j][r]);
    }
    // rows are written as soon as they are complete: a middle dof's row is its output; when a
    // top row is complete, its bottom partner's output is the bottom row plus the previous unit's
    // top row (taken by a shuffle)
    if constexpr (j == kS27 - 1) {
      if constexpr (!RolesW0K1::top(i) && RolesW0K1::partner(i) < 0) {
        *reinterpret_cast<VR *>(zb + static_cast<size_t>(s27_slot(i)) * H * NR) = c[i];
      } else if constexpr (RolesW0K1::top(i)) {
        constexpr int b = s27_bottom_of(i);
        VR tp, out;
#pragma unroll
        for (int r = 0; r < NRL; ++r) tp[r] = __shfl_up(c[i][r], L, 64);
#pragma unroll
        for (int r = 0; r < NRL; ++r) out[r] = c[b][r] + (m > 0 ? tp[r] : T(0));
        *reinterpret_cast<VR *>(zb + static_cast<size_t>(s27_slot(b)) * H * NR) = out;
      }
    }
  });
  // the top layer's unit also writes 0 + each top row at level nl of its bottom partner's column
  if (store && m == nl - 1) {
    s27_for<0, kS27>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      if constexpr (RolesW0K1::top(i)) {
        VR top;
#pragma unroll
        for (int r = 0; r < NRL; ++r) top[r] = T(0) + c[i][r];
        *reinterpret_cast<VR *>(zb + (static_cast<size_t>(s27_slot(s27_bottom_of(i))) * H + 1) * NR) = top;
      }
    });
  }
}

// Phase 2 for point t = col*(nl+1) + m of a problem: false past the end (or at the missing top
// level of a column without one), else acc = y at the point and ydof = the point's dof.
// zu[col]: the column's cells as pos*18 + s27_slot(i) in colour order, -1 past them.
template <typename T, int NR>
__device__ inline bool sym27_gather_point(long long t, const T *__restrict__ z, const int2 *__restrict__ cols,
                                          const int4 *__restrict__ zu, int ncols, int nl,
                                          typename RVec<T, NR>::type &acc, size_t &ydof) {
  typedef typename RVec<T, NR>::type VR;
  const int H = nl + 1;
  if (t >= static_cast<long long>(ncols) * H) return false;
  const int col = static_cast<int>(t / H);
  const int m = static_cast<int>(t - static_cast<long long>(col) * H);
  const int2 ci = cols[col];
  if (m == nl && !ci.y) return false;
  const int4 u4 = zu[col];
  const int us[4] = {u4.x, u4.y, u4.z, u4.w};
  VR v[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    v[q] = us[q] >= 0 ? *reinterpret_cast<const VR *>(z + (static_cast<size_t>(us[q]) * H + m) * NR) : VR{};
  acc = v[0];
#pragma unroll
  for (int q = 1; q < 4; ++q)
    if (us[q] >= 0)
#pragma unroll
      for (int r = 0; r < NR; ++r) acc[r] = acc[r] + v[q][r];
  ydof = static_cast<size_t>(ci.x) + m;
  return true;
}

}  // namespace
}  // namespace lfric
