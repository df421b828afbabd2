/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * LMA operator kernels in the reference's loop order.  The operator is stored as the
 * reference holds it, matrix(ncell_3d, ndf1, ndf2) column-major, with 3D cell index
 * (cell-1)*nlayers + k (infrastructure/source/operator/operator_mod.t90:100,113,119).
 * This file is compiled with -ffp-contract=off so every += is a separate rounding, as in
 * the reference's vector statements.
 */
#include "oracle.h"
#include <string.h>

#define MAT(ik, df, df2, nc3, ndf1) matrix[(ik) + (int64_t)(nc3) * ((df) + (int64_t)(ndf1) * (df2))]

/* invoke(setval_c(y,0), matrix_vector_kernel_type(y,x,A)) on one partition:
 * the PSy layer loops colour by colour over cells_in_colour
 * (applications/lbc_demo/source/psy/sci_psykal_apply_real_lbc_lite_mod.f90:182-203) and
 * calls matrix_vector_code
 * (components/science/source/kernel/algebra/matrix_vector_kernel_mod.F90:107-140):
 *   do df2; i2 = map2(df2); do df; i1 = map1(df)
 *     lhs(i1:i1+nl) = lhs(i1:i1+nl) + matrix(ik:ik+nl,df,df2)*x(i2:i2+nl)            */
void orc_lma_apply(int ncells, int nl, int ncol, const int *npc, const int *cic, int maxcc,
                   int ndf1, int undf1, const int *map1, int ndf2, int undf2, const int *map2,
                   const double *matrix, const double *x, double *y) {
  (void)undf2;
  const int64_t nc3 = (int64_t)ncells * nl;
  memset(y, 0, sizeof(double) * (size_t)undf1);
  for (int colour = 0; colour < ncol; ++colour)
    for (int i = 0; i < npc[colour]; ++i) {
      int cell = cic[colour * maxcc + i];
      const int *m1 = &map1[(size_t)(cell - 1) * ndf1], *m2 = &map2[(size_t)(cell - 1) * ndf2];
      int64_t ik = (int64_t)(cell - 1) * nl;
      for (int df2 = 0; df2 < ndf2; ++df2) {
        int i2 = m2[df2] - 1;
        for (int df = 0; df < ndf1; ++df) {
          int i1 = m1[df] - 1;
          for (int k = 0; k < nl; ++k) y[i1 + k] = y[i1 + k] + MAT(ik + k, df, df2, nc3, ndf1) * x[i2 + k];
        }
      }
    }
}

/* r_single variant (matrix_vector_kernel_mod.F90:68-102): FP32 storage and arithmetic */
void orc_lma_apply_r32(int ncells, int nl, int ncol, const int *npc, const int *cic, int maxcc,
                       int ndf1, int undf1, const int *map1, int ndf2, int undf2,
                       const int *map2, const float *matrix, const float *x, float *y) {
  (void)undf2;
  const int64_t nc3 = (int64_t)ncells * nl;
  memset(y, 0, sizeof(float) * (size_t)undf1);
  for (int colour = 0; colour < ncol; ++colour)
    for (int i = 0; i < npc[colour]; ++i) {
      int cell = cic[colour * maxcc + i];
      const int *m1 = &map1[(size_t)(cell - 1) * ndf1], *m2 = &map2[(size_t)(cell - 1) * ndf2];
      int64_t ik = (int64_t)(cell - 1) * nl;
      for (int df2 = 0; df2 < ndf2; ++df2) {
        int i2 = m2[df2] - 1;
        for (int df = 0; df < ndf1; ++df) {
          int i1 = m1[df] - 1;
          for (int k = 0; k < nl; ++k) y[i1 + k] = y[i1 + k] + MAT(ik + k, df, df2, nc3, ndf1) * x[i2 + k];
        }
      }
    }
}

/* transpose_matrix_vector_code (components/science/source/kernel/algebra/
 * transpose_matrix_vector_kernel_mod.F90): lhs(map1+k) += (A_k^T x_e), matrix is
 * (ncell_3d, ndf2, ndf1); y zeroed first, coloured loop as above. */
void orc_transpose_apply(int ncells, int nl, int ncol, const int *npc, const int *cic,
                         int maxcc, int ndf1, int undf1, const int *map1, int ndf2, int undf2,
                         const int *map2, const double *matrix, const double *x, double *y) {
  (void)undf2;
  const int64_t nc3 = (int64_t)ncells * nl;
  double xe[64], le[64];
  memset(y, 0, sizeof(double) * (size_t)undf1);
  for (int colour = 0; colour < ncol; ++colour)
    for (int i = 0; i < npc[colour]; ++i) {
      int cell = cic[colour * maxcc + i];
      const int *m1 = &map1[(size_t)(cell - 1) * ndf1], *m2 = &map2[(size_t)(cell - 1) * ndf2];
      for (int k = 0; k < nl; ++k) {
        for (int df = 0; df < ndf2; ++df) xe[df] = x[m2[df] - 1 + k];
        int64_t ik = (int64_t)(cell - 1) * nl + k;
        for (int df1 = 0; df1 < ndf1; ++df1) {
          double s = 0.0;
          for (int df = 0; df < ndf2; ++df) s = s + MAT(ik, df, df1, nc3, ndf2) * xe[df];
          le[df1] = s;
        }
        for (int df1 = 0; df1 < ndf1; ++df1) y[m1[df1] - 1 + k] = y[m1[df1] - 1 + k] + le[df1];
      }
    }
}

/* dg_matrix_vector_code (components/science/source/kernel/algebra/
 * dg_matrix_vector_kernel_mod.F90): discontinuous output, zero the column then accumulate;
 * no colouring needed (GH_WRITE). */
void orc_dg_apply(int ncells, int nl, int ndf1, int undf1, const int *map1, int ndf2,
                  int undf2, const int *map2, const double *matrix, const double *x,
                  double *y) {
  (void)undf1; (void)undf2;
  const int64_t nc3 = (int64_t)ncells * nl;
  for (int cell = 1; cell <= ncells; ++cell) {
    const int *m1 = &map1[(size_t)(cell - 1) * ndf1], *m2 = &map2[(size_t)(cell - 1) * ndf2];
    int64_t ik = (int64_t)(cell - 1) * nl;
    for (int df1 = 0; df1 < ndf1; ++df1)
      for (int k = 0; k < nl; ++k) y[m1[df1] - 1 + k] = 0.0;
    for (int df2 = 0; df2 < ndf2; ++df2)
      for (int df1 = 0; df1 < ndf1; ++df1)
        for (int k = 0; k < nl; ++k)
          y[m1[df1] - 1 + k] = y[m1[df1] - 1 + k] + MAT(ik + k, df1, df2, nc3, ndf1) * x[m2[df2] - 1 + k];
  }
}

/* mm_diagonal_kernel_code (components/science/source/kernel/algebra/
 * sci_mm_diagonal_kernel_mod.F90:87-92): diag(map(df)+k) += A(ik,df,df) summed over
 * cells in cell order (diag must be zeroed by the caller, as setval_c does). */
void orc_mm_diagonal(int ncells, int nl, int ndf, int undf, const int *map,
                     const double *matrix, double *diag) {
  (void)undf;
  const int64_t nc3 = (int64_t)ncells * nl;
  for (int cell = 1; cell <= ncells; ++cell) {
    const int *m = &map[(size_t)(cell - 1) * ndf];
    for (int k = 0; k < nl; ++k) {
      int64_t ik = (int64_t)(cell - 1) * nl + k;
      for (int df = 0; df < ndf; ++df) diag[m[df] - 1 + k] = diag[m[df] - 1 + k] + MAT(ik, df, df, nc3, ndf);
    }
  }
}

/* invert_local_operator_code (components/science/source/kernel/algebra/
 * sci_invert_local_operator_kernel_mod.F90:51-77) over every 3D cell, each element inverted by
 * matrix_invert_lu (infrastructure/source/utilities/matrix_invert_mod.F90:62-138 single,
 * :140-216 double): Doolittle LU without pivoting into c (l below the diagonal, u on and above),
 * then column m of the inverse by forward substitution on e_m and back substitution, in the
 * reference's loop order.  Matrices in the reference layout matrix(ncell3d, ndf, ndf). */
#define ORC_INVERT_BODY(T)                                                                     \
  T c[ORC_MAX_NDF][ORC_MAX_NDF], y[ORC_MAX_NDF], x[ORC_MAX_NDF];                               \
  for (int64_t ik = 0; ik < nc3; ++ik) {                                                       \
    for (int i = 0; i < n; ++i)                                                                \
      for (int j = 0; j < n; ++j) c[i][j] = matrix[ik + nc3 * (i + (int64_t)n * j)];           \
    for (int m = 0; m < n - 1; ++m)                                                            \
      for (int i = m + 1; i < n; ++i) {                                                        \
        const T coeff = -(c[i][m] / c[m][m]);                                                  \
        c[i][m] = -coeff; /* l(i,m) */                                                         \
        for (int j = m + 1; j < n; ++j) c[i][j] = c[i][j] + coeff * c[m][j];                   \
      }                                                                                        \
    for (int m = 0; m < n; ++m) {                                                              \
      y[0] = (m == 0) ? (T)1 : (T)0;                                                           \
      for (int i = 1; i < n; ++i) {                                                            \
        y[i] = (i == m) ? (T)1 : (T)0;                                                         \
        for (int j = 0; j < i; ++j) y[i] = y[i] - y[j] * c[i][j];                              \
      }                                                                                        \
      x[n - 1] = y[n - 1] / c[n - 1][n - 1];                                                   \
      for (int i = n - 2; i >= 0; --i) {                                                       \
        x[i] = y[i];                                                                           \
        for (int j = n - 1; j >= i + 1; --j) x[i] = x[i] - x[j] * c[i][j];                     \
        x[i] = x[i] / c[i][i];                                                                 \
      }                                                                                        \
      for (int i = 0; i < n; ++i) inv[ik + nc3 * (i + (int64_t)n * m)] = x[i];                 \
    }                                                                                          \
  }

int orc_invert_local(int64_t nc3, int n, const double *matrix, double *inv) {
  if (n < 1 || n > ORC_MAX_NDF) return 1;
  ORC_INVERT_BODY(double)
  return 0;
}

int orc_invert_local_r32(int64_t nc3, int n, const float *matrix, float *inv) {
  if (n < 1 || n > ORC_MAX_NDF) return 1;
  ORC_INVERT_BODY(float)
  return 0;
}
#undef ORC_INVERT_BODY
