// Multi-RHS LMA apply of lowest-order W0 operators (8 dofs per cell, FP32) on the matrix cores.
//
// Unit of work: one (cell, level m) of a colour, as in the colour-ordered scatter kernels: it
// owns the dofs at level m of the cell's 4 bottom-vertex columns, whose values are the bottom
// rows of the cell's layer-m element plus the top rows of its layer-(m-1) element (both stored
// at level-aligned slot e).  For RHS r and bottom vertex i,
//   c  = sum_j A_m[i][j]     x[b_j + m][r]        (chain of fma over j = 0..7, from 0)
//   cp = sum_j A_m-1[i+4][j] x[b_j + m - 1][r]
//   y[b_i + m][r] (=|+=) c + cp                   (= for the colour that first touches the dof)
// which is the element-matrix x gathered-dofs contraction A(4x8) X(8xNR) of
// matrix_vector_code (components/science/source/kernel/algebra/matrix_vector_kernel_mod.F90:
// 94-100) for the 4 bottom rows.  Each contraction is a chain of 8
// v_mfma_f32_4x4x1f32 steps, one per j: that instruction multiplies 16 independent 4x1 by 1x4
// blocks; a block here is one unit x 4 right-hand sides, lane 4u+s supplying row s of A
// (operand a) and RHS s of X (operand b), and receiving D[v][s] (v = 0..3) in its 4
// accumulator registers.  A K=1 chain rounds exactly as the fma chain of the FMA kernels, so
// the result equals k_lma_colour_mrhs bit for bit.
//
// The caller keeps all 64 lanes of every wavefront in the MFMA steps: lanes without a unit
// pass zero operands and skip the stores.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.h"

namespace lfric {
namespace {

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// row r (8 entries) of element slot e of an FP32 operator in the chunked AoSoA layout
__device__ inline void load_row8(const float *__restrict__ A, size_t gs, long long e, int r, float (&a)[8]) {
  for_row<float, 8>(A, gs, e, r, [&](int j, float v) { a[j] = v; });
}

// NR right-hand sides (1, 2, 4 or 8), in groups of 4.  A W0 cell's top vertex j+4 lies one
// level above its bottom vertex j in the same column, so the x entries of a unit are its 4
// bottom columns at levels m-1, m and m+1.  Each lane loads level m for its RHS; the unit in
// the previous 4 lanes (same cell, level m-1) and the one in the next 4 lanes (level m+1)
// hold the other two, which are taken by lane shuffles (own loads at a wavefront or column
// edge).
template <int NR>
__device__ inline void mfma_unit_w0(const float *__restrict__ A, size_t gs, long long e, const int *__restrict__ map1p,
                                    const int *__restrict__ map2p, uint32_t first, const float *__restrict__ x,
                                    float *__restrict__ y, int m, int nl, bool valid) {
  const int lane = threadIdx.x & 63, s = lane & 3;
  const bool has_m = valid && m < nl, has_p = valid && m > 0;
  // whether the neighbouring units (lane -/+ 4) are this cell's levels m-1 / m+1
  const bool up_ok = has_p && lane >= 4, dn_ok = has_m && lane < 60;
  float am[8], ap[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) am[j] = ap[j] = 0.f;
  if (has_m) load_row8(A, gs, e, s, am);
  if (has_p) load_row8(A, gs, e, s + 4, ap);
  int b[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = valid ? map2p[j] : 0;
  constexpr int NG = (NR + 3) / 4;
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const int r = g * 4 + s;
    const bool rv = r < NR;
    float x0[4], xm1[4], xp1[4];  // levels m, m-1, m+1 of the 4 bottom columns, RHS r
#pragma unroll
    for (int j = 0; j < 4; ++j) x0[j] = (valid && rv) ? x[(static_cast<size_t>(b[j]) + m) * NR + r] : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float up = __shfl_up(x0[j], 4, 64), dn = __shfl_down(x0[j], 4, 64);
      xm1[j] = up_ok ? up : (has_p && rv) ? x[(static_cast<size_t>(b[j]) + m - 1) * NR + r] : 0.f;
      xp1[j] = dn_ok ? dn : (has_m && rv) ? x[(static_cast<size_t>(b[j]) + m + 1) * NR + r] : 0.f;
    }
    mfma_f4 c = {0.f, 0.f, 0.f, 0.f}, cp = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // bottom rows of layer m: x at levels m (j < 4) and m+1
      const float xv = (has_m && rv) ? (j < 4 ? x0[j] : xp1[j - 4]) : 0.f;
      c = __builtin_amdgcn_mfma_f32_4x4x1f32(am[j], xv, c, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // top rows of layer m-1: x at levels m-1 (j < 4) and m
      const float xv = (has_p && rv) ? (j < 4 ? xm1[j] : x0[j - 4]) : 0.f;
      cp = __builtin_amdgcn_mfma_f32_4x4x1f32(ap[j], xv, cp, 0, 0, 0);
    }
    if (valid && rv) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float *yd = y + (static_cast<size_t>(map1p[v]) + m) * NR + r;
        const float out = c[v] + cp[v];
        if ((first >> v) & 1u)
          *yd = out;
        else
          *yd = *yd + out;
      }
    }
  }
}

}  // namespace
}  // namespace lfric
