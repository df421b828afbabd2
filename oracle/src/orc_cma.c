/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * Columnwise (CMA) operators: banded per-column matrices of a horizontally discontinuous
 * function space, and the vertical tridiagonal solve.  Matrices are in the reference layout
 * columnwise_matrix(bandwidth, nrow, ncell_2d), Fortran order; p[7] = {alpha, beta, gamma_m,
 * gamma_p, bandwidth, nrow, ncol}.  Every loop keeps the reference's 1-based indices so the
 * restatement can be read against the Fortran line by line; float and double instances are
 * generated from one body.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

#define CMA(M, P, b, i, cell) (M)[((b) - 1) + (int64_t)(P)[4] * (((i) - 1) + (int64_t)(P)[5] * ((cell) - 1))]

static int orc_gcd(int a, int b) {
  /* get_gcd (infrastructure/source/operator/columnwise_operator_mod.f90): Euclid on |a|, |b| */
  int x = abs(a), y = abs(b);
  while (y != 0) {
    const int t = x % y;
    x = y;
    y = t;
  }
  return x;
}

/* columnwise_operator_initialiser / _custom and divide_by_gcd
 * (infrastructure/source/operator/columnwise_operator_mod.f90:89-130,175-190, extract_mesh_fs_info
 * :270-282): nrow = nl*ni_to + (nl+1)*nf_to, ncol likewise; alpha = ni_from + nf_from, beta =
 * ni_to + nf_to, gamma_m = (ndf_from - 1)*beta, gamma_p = (ndf_to - 1)*alpha unless given, all
 * four divided by gcd(alpha, beta); bandwidth = 1 + ceiling((gamma_m + gamma_p)/beta) in
 * r_solver arithmetic.  custom = 0 uses the derived alpha..gamma_p. */
int orc_cma_params(int nl, int ni_to, int nf_to, int ni_from, int nf_from, int custom, int alpha, int beta,
                   int gamma_m, int gamma_p, int *p) {
  const int ndf_to = ni_to + 2 * nf_to, ndf_from = ni_from + 2 * nf_from;
  if (!custom) {
    alpha = ni_from + nf_from;
    beta = ni_to + nf_to;
    gamma_m = (ndf_from - 1) * (ni_to + nf_to);
    gamma_p = (ndf_to - 1) * (ni_from + nf_from);
  }
  const int g = orc_gcd(alpha, beta);
  if (g == 0) return 1;
  if (g > 1) {
    alpha /= g;
    beta /= g;
    gamma_m /= g;
    gamma_p /= g;
  }
  p[0] = alpha;
  p[1] = beta;
  p[2] = gamma_m;
  p[3] = gamma_p;
  p[4] = 1 + (int)ceilf((float)(gamma_m + gamma_p) / (float)beta);
  p[5] = nl * ni_to + (nl + 1) * nf_to;
  p[6] = nl * ni_from + (nl + 1) * nf_from;
  return 0;
}

/* build_dofmaps (columnwise_operator_mod.f90:196-230): from cell 1's dofmap, column_dofmap(df,k)
 * = map(df) - map(1) + 1 + (k-1), column_banded_dofmap(df,k) = (k-1)*(ni + nf) + df and
 * indirection(banded(df,k)) = column_dofmap(df,k).  banded is (ndf, nl), indirection (nrow). */
void orc_cma_dofmaps(int nl, int ndf, int ni, int nf, const int *map_cell1, int *banded, int *indirection) {
  for (int k = 1; k <= nl; ++k)
    for (int df = 1; df <= ndf; ++df) {
      const int col = (map_cell1[df - 1] - map_cell1[0] + 1) + (k - 1);
      const int b = (k - 1) * (ni + nf) + df;
      banded[(df - 1) + ndf * (k - 1)] = b;
      indirection[b - 1] = col;
    }
}

#define ORC_CMA_BODIES(T, SFX)                                                                              \
  static int jminus##SFX(const int *p, int i) {                                                             \
    return (int)ceil((double)((T)(p[0] * i - p[3]) / ((T)1 * (T)p[1])));                                    \
  }                                                                                                         \
  static int jplus##SFX(const int *p, int i) {                                                              \
    return (int)floor((double)((T)(p[0] * i + p[2]) / ((T)1 * (T)p[1])));                                   \
  }                                                                                                         \
  /* columnwise_op_asm_kernel_code (components/science/source/kernel/columnwise_operators/                 \
   * sci_columnwise_op_asm_kernel_mod.F90:96-140): zero the column, then add local_stencil(ik,df1,df2)      \
   * at band j - j_minus + 1 of row i, looping k, df1, df2. */                                              \
  void orc_cma_asm##SFX(int ncells, int nl, int ndf_to, int ndf_from, const T *lma, const int *p,           \
                        const int *banded_to, const int *banded_from, T *cma) {                             \
    const int64_t nc3 = (int64_t)ncells * nl;                                                               \
    for (int cell = 1; cell <= ncells; ++cell) {                                                            \
      for (int i = 1; i <= p[5]; ++i)                                                                       \
        for (int b = 1; b <= p[4]; ++b) CMA(cma, p, b, i, cell) = (T)0;                                     \
      for (int k = 1; k <= nl; ++k) {                                                                       \
        const int64_t ik = (int64_t)(cell - 1) * nl + k;                                                    \
        for (int df1 = 1; df1 <= ndf_to; ++df1) {                                                           \
          const int i = banded_to[(df1 - 1) + ndf_to * (k - 1)];                                            \
          const int jm = jminus##SFX(p, i);                                                                 \
          for (int df2 = 1; df2 <= ndf_from; ++df2) {                                                       \
            const int j = banded_from[(df2 - 1) + ndf_from * (k - 1)];                                      \
            CMA(cma, p, j - jm + 1, i, cell) =                                                              \
                CMA(cma, p, j - jm + 1, i, cell) + lma[(ik - 1) + nc3 * ((df1 - 1) + (int64_t)ndf_to * (df2 - 1))]; \
          }                                                                                                 \
        }                                                                                                   \
      }                                                                                                     \
    }                                                                                                       \
  }                                                                                                         \
  /* columnwise_op_app_kernel_code (sci_columnwise_op_app_kernel_mod.F90:107-131): for each row i,        \
   * lhs(mu_i) += sum over j in [max(1, j_minus), min(ncol, j_plus)] of A(j - j_minus + 1, i) x(mu_j),     \
   * mu = map(1) + indirection - 1.  lhs accumulates (the caller zeroes it, as setval_c does). */          \
  void orc_cma_apply##SFX(int ncells, const int *p, const T *cma, const T *x, T *lhs, int ndf1,            \
                          const int *map1, const int *ind_to, int ndf2, const int *map2,                    \
                          const int *ind_from) {                                                            \
    for (int cell = 1; cell <= ncells; ++cell)                                                              \
      for (int i = 1; i <= p[5]; ++i) {                                                                     \
        const int64_t mu_i = map1[(int64_t)ndf1 * (cell - 1)] + ind_to[i - 1] - 1;                          \
        const int jm = jminus##SFX(p, i), jp = jplus##SFX(p, i);                                            \
        for (int j = jm > 1 ? jm : 1; j <= (p[6] < jp ? p[6] : jp); ++j) {                                  \
          const int64_t mu_j = map2[(int64_t)ndf2 * (cell - 1)] + ind_from[j - 1] - 1;                      \
          lhs[mu_i - 1] = lhs[mu_i - 1] + CMA(cma, p, j - jm + 1, i, cell) * x[mu_j - 1];                   \
        }                                                                                                   \
      }                                                                                                     \
  }                                                                                                         \
  /* columnwise_op_appinv_kernel_code (sci_columnwise_op_appinv_kernel_mod.F90:107-146): Thomas         \
   * algorithm on the tridiagonal column (bands 1, 2, 3 = sub, diagonal, super). */                        \
  void orc_cma_appinv##SFX(int ncells, const int *p, const T *cma, const T *x, T *lhs, int ndf,            \
                           const int *map, const int *ind) {                                                \
    const int nrow = p[5];                                                                                  \
    T *cp = (T *)malloc(sizeof(T) * (size_t)nrow), *dp = (T *)malloc(sizeof(T) * (size_t)nrow);             \
    for (int cell = 1; cell <= ncells; ++cell) {                                                            \
      const int64_t base = map[(int64_t)ndf * (cell - 1)];                                                  \
      for (int i = 1; i <= nrow; ++i) {                                                                     \
        const int64_t mu_i = base + ind[i - 1] - 1;                                                         \
        if (i == 1) {                                                                                       \
          const T inv = (T)1 / CMA(cma, p, 2, i, cell);                                                     \
          cp[i - 1] = inv * CMA(cma, p, 3, i, cell);                                                        \
          dp[i - 1] = inv * x[mu_i - 1];                                                                    \
        } else {                                                                                            \
          const T inv = (T)1 / (CMA(cma, p, 2, i, cell) - CMA(cma, p, 1, i, cell) * cp[i - 2]);             \
          if (i < nrow) cp[i - 1] = inv * CMA(cma, p, 3, i, cell);                                          \
          dp[i - 1] = inv * (x[mu_i - 1] - CMA(cma, p, 1, i, cell) * dp[i - 2]);                            \
        }                                                                                                   \
      }                                                                                                     \
      for (int i = nrow; i >= 1; --i) {                                                                     \
        const int64_t mu_i = base + ind[i - 1] - 1;                                                         \
        if (i < nrow) dp[i - 1] = dp[i - 1] - cp[i - 1] * dp[i];                                            \
        lhs[mu_i - 1] = dp[i - 1];                                                                          \
      }                                                                                                     \
    }                                                                                                       \
    free(cp);                                                                                               \
    free(dp);                                                                                               \
  }                                                                                                         \
  /* columnwise_op_mul_kernel_code (sci_columnwise_op_mul_kernel_mod.F90:128-160): C += A B, each        \
   * entry of C's band summed over ell in A's row range where B's row ell reaches column j. */              \
  void orc_cma_mul##SFX(int ncells, const int *pa, const T *A, const int *pb, const T *B, const int *pc,    \
                        T *C) {                                                                             \
    for (int cell = 1; cell <= ncells; ++cell)                                                              \
      for (int i = 1; i <= pc[5]; ++i) {                                                                    \
        const int jma = jminus##SFX(pa, i), jpa = jplus##SFX(pa, i);                                        \
        const int jmc = jminus##SFX(pc, i), jpc = jplus##SFX(pc, i);                                        \
        for (int j = jmc > 1 ? jmc : 1; j <= (pc[6] < jpc ? pc[6] : jpc); ++j) {                            \
          T e = (T)0;                                                                                       \
          for (int l = jma > 1 ? jma : 1; l <= (pa[6] < jpa ? pa[6] : jpa); ++l) {                          \
            const int jmb = jminus##SFX(pb, l), jpb = jplus##SFX(pb, l);                                    \
            if (jmb <= j && j <= jpb) e = e + CMA(A, pa, l - jma + 1, i, cell) * CMA(B, pb, j - jmb + 1, l, cell); \
          }                                                                                                 \
          CMA(C, pc, j - jmc + 1, i, cell) = CMA(C, pc, j - jmc + 1, i, cell) + e;                          \
        }                                                                                                   \
      }                                                                                                     \
  }                                                                                                         \
  /* columnwise_op_scaledadd_kernel_code (sci_columnwise_op_scaledadd_kernel_mod.F90:132-166):           \
   * C = 0, then alpha A then beta B added band by band (B's loop bounded by A's ncol, as there). */        \
  void orc_cma_scaledadd##SFX(int ncells, const int *pa, const T *A, const int *pb, const T *B,            \
                              const int *pc, T *C, T alpha, T beta) {                                       \
    for (int cell = 1; cell <= ncells; ++cell) {                                                            \
      for (int i = 1; i <= pc[5]; ++i)                                                                      \
        for (int b = 1; b <= pc[4]; ++b) CMA(C, pc, b, i, cell) = (T)0;                                     \
      for (int i = 1; i <= pa[5]; ++i) {                                                                    \
        const int jma = jminus##SFX(pa, i), jpa = jplus##SFX(pa, i);                                        \
        const int jmb = jminus##SFX(pb, i), jpb = jplus##SFX(pb, i);                                        \
        const int jmc = jminus##SFX(pc, i);                                                                 \
        for (int j = jma > 1 ? jma : 1; j <= (pa[6] < jpa ? pa[6] : jpa); ++j)                              \
          CMA(C, pc, j - jmc + 1, i, cell) = CMA(C, pc, j - jmc + 1, i, cell) + alpha * CMA(A, pa, j - jma + 1, i, cell); \
        for (int j = jmb > 1 ? jmb : 1; j <= (pa[6] < jpb ? pa[6] : jpb); ++j)                              \
          CMA(C, pc, j - jmc + 1, i, cell) = CMA(C, pc, j - jmc + 1, i, cell) + beta * CMA(B, pb, j - jmb + 1, i, cell); \
      }                                                                                                     \
    }                                                                                                       \
  }                                                                                                         \
  /* columnwise_op_asm_diag_hmht_kernel_code (sci_columnwise_op_asm_diag_hmht_kernel_mod.F90:95-135):    \
   * the W3 column band of Dh diag(M2h)^-1 Dh^T, tmp summed over df3 of Dh(df1,df3)*Dh(df2,df3)/M2h(df3,df3). */ \
  void orc_cma_asm_diag_hmht##SFX(int ncells, int nl, int ndf_w3, int ndf_w2h, const T *Dh, const T *M2h,  \
                                  const int *p, const int *banded, T *cma) {                                \
    const int64_t nc3 = (int64_t)ncells * nl;                                                               \
    for (int cell = 1; cell <= ncells; ++cell) {                                                            \
      for (int i = 1; i <= p[5]; ++i)                                                                       \
        for (int b = 1; b <= p[4]; ++b) CMA(cma, p, b, i, cell) = (T)0;                                     \
      for (int k = 1; k <= nl; ++k) {                                                                       \
        const int64_t ik = (int64_t)(cell - 1) * nl + k;                                                    \
        for (int df1 = 1; df1 <= ndf_w3; ++df1) {                                                           \
          const int i = banded[(df1 - 1) + ndf_w3 * (k - 1)];                                               \
          const int jm = jminus##SFX(p, i);                                                                 \
          for (int df2 = 1; df2 <= ndf_w3; ++df2) {                                                         \
            T tmp = (T)0;                                                                                   \
            for (int df3 = 1; df3 <= ndf_w2h; ++df3)                                                        \
              tmp = tmp + Dh[(ik - 1) + nc3 * ((df1 - 1) + (int64_t)ndf_w3 * (df3 - 1))] *                  \
                              Dh[(ik - 1) + nc3 * ((df2 - 1) + (int64_t)ndf_w3 * (df3 - 1))] /              \
                              M2h[(ik - 1) + nc3 * ((df3 - 1) + (int64_t)ndf_w2h * (df3 - 1))];             \
            const int j = banded[(df2 - 1) + ndf_w3 * (k - 1)];                                             \
            CMA(cma, p, j - jm + 1, i, cell) = CMA(cma, p, j - jm + 1, i, cell) + tmp;                      \
          }                                                                                                 \
        }                                                                                                   \
      }                                                                                                     \
    }                                                                                                       \
  }                                                                                                         \
  /* columnwise_op_asm_m2v_lumped_inv_krnl_code (sci_columnwise_op_asm_m2v_lumped_inv_krnl_mod.F90:       \
   * 85-113): band 1 of row i sums the diagonal entries local_stencil(ik,df,df) that land on it, then     \
   * every nonzero entry of the column is replaced by its reciprocal. */                                    \
  void orc_cma_asm_m2v_lumped_inv##SFX(int ncells, int nl, int ndf, const T *lma, const int *p,            \
                                       const int *banded, T *cma) {                                         \
    const int64_t nc3 = (int64_t)ncells * nl;                                                               \
    for (int cell = 1; cell <= ncells; ++cell) {                                                            \
      for (int i = 1; i <= p[5]; ++i)                                                                       \
        for (int b = 1; b <= p[4]; ++b) CMA(cma, p, b, i, cell) = (T)0;                                     \
      for (int k = 1; k <= nl; ++k) {                                                                       \
        const int64_t ik = (int64_t)(cell - 1) * nl + k;                                                    \
        for (int df = 1; df <= ndf; ++df) {                                                                 \
          const int i = banded[(df - 1) + ndf * (k - 1)];                                                   \
          CMA(cma, p, 1, i, cell) = CMA(cma, p, 1, i, cell) + lma[(ik - 1) + nc3 * ((df - 1) + (int64_t)ndf * (df - 1))]; \
        }                                                                                                   \
      }                                                                                                     \
      for (int i = 1; i <= p[5]; ++i)                                                                       \
        for (int b = 1; b <= p[4]; ++b)                                                                     \
          if (CMA(cma, p, b, i, cell) != (T)0) CMA(cma, p, b, i, cell) = (T)1 / CMA(cma, p, b, i, cell);    \
    }                                                                                                       \
  }                                                                                                         \
  /* tri_solve_code (components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90:57-130): the    \
   * W3 column tridiagonal system with diagonal tri_0, super-diagonal tri_plus, sub-diagonal tri_minus,    \
   * forward elimination then back substitution.  map is the W3 dofmap (1 dof per cell). */                \
  void orc_tri_solve##SFX(int ncells, int nl, const int *map, const T *x, const T *t0, const T *tp,        \
                          const T *tm, T *y) {                                                              \
    T *tpn = (T *)malloc(sizeof(T) * (size_t)nl), *xn = (T *)malloc(sizeof(T) * (size_t)nl);                \
    for (int cell = 0; cell < ncells; ++cell) {                                                             \
      const int64_t ij = map[cell] - 1;                                                                     \
      T denom = (T)1 / t0[ij];                                                                              \
      tpn[0] = tp[ij] * denom;                                                                              \
      xn[0] = x[ij] * denom;                                                                                \
      for (int k = 1; k <= nl - 1; ++k) {                                                                   \
        denom = (T)1 / (t0[ij + k] - tm[ij + k] * tpn[k - 1]);                                              \
        tpn[k] = tp[ij + k] * denom;                                                                        \
        xn[k] = (x[ij + k] - tm[ij + k] * xn[k - 1]) * denom;                                               \
      }                                                                                                     \
      y[ij + nl - 1] = xn[nl - 1];                                                                          \
      for (int k = nl - 2; k >= 0; --k) y[ij + k] = xn[k] - tpn[k] * y[ij + k + 1];                         \
    }                                                                                                       \
    free(tpn);                                                                                              \
    free(xn);                                                                                               \
  }

ORC_CMA_BODIES(double, )
ORC_CMA_BODIES(float, _r32)
