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
#include <math.h>
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

/* dg_inc_matrix_vector_code (components/science/source/kernel/algebra/
 * dg_inc_matrix_vector_kernel_mod.F90:69-104, real64 variant :107-142): lhs is GH_READWRITE on a
 * discontinuous space, so the cells are visited in order with no colouring and nothing is
 * zeroed: lhs(i1:i1+nl) = lhs(i1:i1+nl) + matrix(ik:ik+nl,df,df2)*x(i2:i2+nl), df2 outer. */
void orc_dg_inc_apply(int ncells, int nl, int ndf1, int undf1, const int *map1, int ndf2,
                      int undf2, const int *map2, const double *matrix, const double *x,
                      double *y) {
  (void)undf1; (void)undf2;
  const int64_t nc3 = (int64_t)ncells * nl;
  for (int cell = 1; cell <= ncells; ++cell) {
    const int *m1 = &map1[(size_t)(cell - 1) * ndf1], *m2 = &map2[(size_t)(cell - 1) * ndf2];
    const int64_t ik = (int64_t)(cell - 1) * nl;
    for (int df2 = 0; df2 < ndf2; ++df2)
      for (int df1 = 0; df1 < ndf1; ++df1)
        for (int k = 0; k < nl; ++k)
          y[m1[df1] - 1 + k] = y[m1[df1] - 1 + k] + MAT(ik + k, df1, df2, nc3, ndf1) * x[m2[df2] - 1 + k];
  }
}

/* enforce_bc_code (components/science/source/kernel/algebra/sci_enforce_bc_kernel_mod.F90:63-91,
 * real64 variant) over every cell column in cell order: the bottom level of each dof is scaled
 * by boundary_value(df,1), the top layer's (k = nlayers-1) by boundary_value(df,2).
 * bv is boundary_value(ndf,2) in Fortran order (bv[df] bottom, bv[ndf+df] top). */
void orc_enforce_bc(int ncells, int nl, int ndf, const int *map, const int *bv, double *field) {
  for (int cell = 1; cell <= ncells; ++cell) {
    const int *m = &map[(size_t)(cell - 1) * ndf];
    for (int df = 0; df < ndf; ++df) field[m[df] - 1] = field[m[df] - 1] * (double)bv[df];
    for (int df = 0; df < ndf; ++df)
      field[m[df] - 1 + nl - 1] = field[m[df] - 1 + nl - 1] * (double)bv[ndf + df];
  }
}

/* dof_on_vert_boundary of a lowest-order space
 * (infrastructure/source/function_space/function_space_constructor_helper_functions_mod.F90:
 * 678-680: every flag 1, then per space): W2 (:1030-1049) flags its face dofs on face
 * number_faces-1 (bottom, dof 5) and number_faces (top, dof 6); W0 (:690-800) and W3 (:1172-)
 * flag none.  space: 0 W0, 2 W2, 3 W3.  Returns 0, or -1 for a space/order not restated. */
int orc_boundary_dofs(int space, int order, int ndf, int *bv) {
  for (int i = 0; i < 2 * ndf; ++i) bv[i] = 1;
  if (space == 0 || space == 3) return 0;
  if (space == 2 && order == 0 && ndf == 6) {
    bv[4] = 0;           /* face B: boundary_value(5,1) = 0 */
    bv[ndf + 5] = 0;     /* face T: boundary_value(6,2) = 0 */
    return 0;
  }
  return -1;
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

/* Gauss-Legendre rule of n points on [0, 1]: quadrature_rule of
 * infrastructure/source/quadrature/quadrature_rule_gaussian_mod.F90:35-94 (Newton iteration on
 * the Legendre recurrence from cos(pi (i - 1/4) / (n + 1/2)) until |z - z1| <= epsilon, z1
 * carried over from the previous root, then the map from [-1, 1]).  The restatement caps the
 * iteration at 100 steps. */
int orc_gauss_rule(int n, double *pts, double *wts) {
  if (n < 1 || n > 16) return 1;
  const double pi = 4.0 * atan(1.0), eps = 2.220446049250313e-16;
  const int m = (n + 1) / 2;
  double z1 = 0.0;  // set once: each root's iteration starts from the previous root
  for (int i = 1; i <= m; ++i) {
    double z = cos(pi * (i - 0.25) / (n + 0.5)), pp = 1.0;
    for (int it = 0; it < 100 && fabs(z - z1) > eps; ++it) {
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
    pts[n - i] = +z;
    wts[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
    wts[n - i] = wts[i - 1];
  }
  for (int i = 0; i < n; ++i) {
    pts[i] = 0.5 * (pts[i] + 1.0);
    wts[i] = 0.5 * wts[i];
  }
  return 0;
}

/* Quadrature and W0 (coordinate) basis tables shared by the assembly restatements.  Gauss
 * points from orc_gauss_rule; horizontal points pair x_i with y_j, y fastest
 * (infrastructure/source/quadrature/quadrature_xyoz_mod.F90:235-243).  Trilinear basis on the
 * reference cube, vertices SWB, SEB, NEB, NWB, SWT, SET, NET, NWT (reference_element_mod.F90:
 * 1244-1251): N = lx*ly*lz with l = t or 1 - t, derivatives with +-1 in place of the
 * differentiated factor. */
enum { ORC_NQMAX = 8 };
typedef struct {
  int nq, nh;
  double p[ORC_NQMAX], w[ORC_NQMAX], wh[ORC_NQMAX * ORC_NQMAX];
  double b[8][ORC_NQMAX * ORC_NQMAX][ORC_NQMAX], db[3][8][ORC_NQMAX * ORC_NQMAX][ORC_NQMAX];
} orc_tab;

static const int orc_vx[8] = {0, 1, 1, 0, 0, 1, 1, 0}, orc_vy[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                 orc_vz[8] = {0, 0, 0, 0, 1, 1, 1, 1};

static int orc_tab_init(orc_tab *t, int nqp) {
  if (nqp < 1 || nqp > ORC_NQMAX) return 1;
  t->nq = nqp;
  t->nh = nqp * nqp;
  orc_gauss_rule(nqp, t->p, t->w);
  for (int i = 0; i < nqp; ++i)
    for (int j = 0; j < nqp; ++j) {
      const int ic = i * nqp + j;
      t->wh[ic] = t->w[i] * t->w[j];
      for (int kz = 0; kz < nqp; ++kz)
        for (int df = 0; df < 8; ++df) {
          const double x = t->p[i], y = t->p[j], z = t->p[kz];
          const double lx = orc_vx[df] ? x : 1.0 - x, ly = orc_vy[df] ? y : 1.0 - y, lz = orc_vz[df] ? z : 1.0 - z;
          const double dx = orc_vx[df] ? 1.0 : -1.0, dy = orc_vy[df] ? 1.0 : -1.0, dz = orc_vz[df] ? 1.0 : -1.0;
          t->b[df][ic][kz] = lx * ly * lz;
          t->db[0][df][ic][kz] = dx * ly * lz;
          t->db[1][df][ic][kz] = lx * dy * lz;
          t->db[2][df][ic][kz] = lx * ly * dz;
        }
    }
  return 0;
}

/* coordinate_jacobian for coord_system_xyz
 * (components/science/source/kernel/geometry/sci_coordinate_jacobian_mod.F90:289-445): jac(i,dir)
 * = sum_df chi_i(df) d_dir N_df in df order, dj by cofactors along the first row.  c[3][8] are
 * the cell's vertex coordinates. */
static double orc_jacobian(const orc_tab *t, const double c[3][8], int ic, int kz, double J[3][3]) {
  for (int r = 0; r < 3; ++r)
    for (int dir = 0; dir < 3; ++dir) J[r][dir] = 0.0;
  for (int df = 0; df < 8; ++df)
    for (int dir = 0; dir < 3; ++dir)
      for (int r = 0; r < 3; ++r) J[r][dir] = J[r][dir] + c[r][df] * t->db[dir][df][ic][kz];
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

static void orc_cell_coords(const int *mc, int k, const double *chi1, const double *chi2, const double *chi3,
                            double c[3][8]) {
  for (int df = 0; df < 8; ++df) {
    c[0][df] = chi1[mc[df] - 1 + k];
    c[1][df] = chi2[mc[df] - 1 + k];
    c[2][df] = chi3[mc[df] - 1 + k];
  }
}

/* Lowest-order W0 mass matrix from Cartesian vertex coordinates: compute_mass_matrix_w_scalar_code
 * (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w_scalar_mod.F90:290-384,
 * r_double) with the Jacobian above.  Entries df >= df2 are summed over (qp_v outer, qp_h inner)
 * of wh*wv*N_df*N_df2*dj and mirrored.  chi1..3 are W0 fields (undf); map is the W0 dofmap
 * (1-based, [cell][8]); mm in the reference layout (ncell_3d, 8, 8). */
int orc_mass_matrix_w0(int ncells, int nl, const int *map, const double *chi1, const double *chi2,
                       const double *chi3, int nqp, double *mm) {
  static orc_tab t;
  if (orc_tab_init(&t, nqp)) return 1;
  const int nh = t.nh, nv = nqp;
  const int64_t nc3 = (int64_t)ncells * nl;
  for (int cell = 0; cell < ncells; ++cell) {
    for (int k = 0; k < nl; ++k) {
      const int64_t ik = (int64_t)cell * nl + k;
      double c[3][8], dj[ORC_NQMAX * ORC_NQMAX][ORC_NQMAX], J[3][3];
      orc_cell_coords(&map[(size_t)cell * 8], k, chi1, chi2, chi3, c);
      for (int kz = 0; kz < nv; ++kz)
        for (int ic = 0; ic < nh; ++ic) dj[ic][kz] = orc_jacobian(&t, c, ic, kz, J);
      for (int df2 = 0; df2 < 8; ++df2) {
        for (int df = df2; df < 8; ++df) {
          double s = 0.0;
          for (int kz = 0; kz < nv; ++kz)
            for (int ic = 0; ic < nh; ++ic)
              s = s + t.wh[ic] * t.w[kz] * t.b[df][ic][kz] * t.b[df2][ic][kz] * dj[ic][kz];
          mm[ik + nc3 * (df + 8 * (int64_t)df2)] = s;
        }
        for (int df = df2 - 1; df >= 0; --df) mm[ik + nc3 * (df + 8 * (int64_t)df2)] = mm[ik + nc3 * (df2 + 8 * (int64_t)df)];
      }
    }
  }
  return 0;
}

/* Lowest-order W2 basis (Raviart-Thomas on the reference cube) as basis_setup builds it
 * (infrastructure/source/function_space/function_space_constructor_helper_functions_mod.F90:
 * 973-1070 with k_h = k_v = 0: one dof per face W, S, E, N, B, T, unit vector the face's
 * normals_to_faces entry, reference_element_mod.F90:1436-1441, i.e. +i, -j, +i, -j, +k, +k,
 * times the linear Lagrange factor that is 1 on that face): each dof has one nonzero
 * component, comp[d], of value sign[d] * (1 - t) on W, S, B and sign[d] * t on E, N, T, where t
 * is x, y or z; its divergence is the constant ddiv[d]. */
static const int orc_w2_comp[6] = {0, 1, 0, 1, 2, 2};
static const double orc_w2_sign[6] = {1.0, -1.0, 1.0, -1.0, 1.0, 1.0};
static const int orc_w2_far[6] = {0, 0, 1, 1, 0, 1};  /* factor t (1) or 1 - t (0) */
static const double orc_w2_div[6] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0};

static double orc_w2_value(const orc_tab *t, int d, int ic, int kz) {
  const int i = ic / t->nq, j = ic % t->nq;
  const double xyz[3] = {t->p[i], t->p[j], t->p[kz]};
  const double s = xyz[orc_w2_comp[d]];
  return orc_w2_sign[d] * (orc_w2_far[d] ? s : 1.0 - s);
}

/* basis_w2(3, 6, nqp_h, nqp_v) and diff_basis_w2(1, 6, nqp_h, nqp_v) in the reference layout. */
int orc_w2_basis(int nqp, double *basis, double *div) {
  static orc_tab t;
  if (orc_tab_init(&t, nqp)) return 1;
  for (int kz = 0; kz < nqp; ++kz)
    for (int ic = 0; ic < t.nh; ++ic)
      for (int d = 0; d < 6; ++d) {
        const size_t q = (size_t)ic + (size_t)t.nh * kz;
        for (int r = 0; r < 3; ++r) basis[r + 3 * (d + 6 * q)] = 0.0;
        basis[orc_w2_comp[d] + 3 * (d + 6 * q)] = orc_w2_value(&t, d, ic, kz);
        div[d + 6 * q] = orc_w2_div[d];
      }
  return 0;
}

/* Lowest-order W2 mass matrix: compute_mass_matrix_w2_code
 * (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w2_mod.F90:83-164).
 * Entries df >= df2 are sums over (qp_v outer, qp_h inner) of
 * wh*wv*dot_product(matmul(jac, b_df), matmul(jac, b_df2))/dj, mirrored.  matmul of jac with a
 * basis vector whose one nonzero component is v in column c is v*jac(:, c) (the other two
 * products are zeros, which leave every nonzero sum unchanged); dot_product sums the three
 * component products in order.  mm in the reference layout (ncell_3d, 6, 6). */
int orc_mass_matrix_w2(int ncells, int nl, const int *map, const double *chi1, const double *chi2,
                       const double *chi3, int nqp, double *mm) {
  static orc_tab t;
  if (orc_tab_init(&t, nqp)) return 1;
  const int nh = t.nh, nv = nqp;
  const int64_t nc3 = (int64_t)ncells * nl;
  for (int cell = 0; cell < ncells; ++cell) {
    for (int k = 0; k < nl; ++k) {
      const int64_t ik = (int64_t)cell * nl + k;
      double c[3][8], s[6][6] = {{0}};
      orc_cell_coords(&map[(size_t)cell * 8], k, chi1, chi2, chi3, c);
      for (int kz = 0; kz < nv; ++kz)
        for (int ic = 0; ic < nh; ++ic) {
          double J[3][3], jb[6][3];
          const double dj = orc_jacobian(&t, c, ic, kz, J);
          for (int d = 0; d < 6; ++d) {
            const double v = orc_w2_value(&t, d, ic, kz);
            for (int r = 0; r < 3; ++r) jb[d][r] = J[r][orc_w2_comp[d]] * v;
          }
          for (int df2 = 0; df2 < 6; ++df2)
            for (int df = df2; df < 6; ++df) {
              const double dot = jb[df][0] * jb[df2][0] + jb[df][1] * jb[df2][1] + jb[df][2] * jb[df2][2];
              s[df][df2] = s[df][df2] + t.wh[ic] * t.w[kz] * dot / dj;
            }
        }
      for (int df2 = 0; df2 < 6; ++df2) {
        for (int df = df2; df < 6; ++df) mm[ik + nc3 * (df + 6 * (int64_t)df2)] = s[df][df2];
        for (int df = df2 - 1; df >= 0; --df) mm[ik + nc3 * (df + 6 * (int64_t)df2)] = s[df2][df];
      }
    }
  }
  return 0;
}

/* Lowest-order Wtheta mass matrix: compute_mass_matrix_w_scalar_code
 * (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w_scalar_mod.F90:290-384)
 * with the Wtheta basis (bottom dof 1 - z, top dof z, constant in x and y: the W2 B and T factors
 * above).  mm in the reference layout (ncell_3d, 2, 2). */
int orc_mass_matrix_wtheta(int ncells, int nl, const int *map, const double *chi1, const double *chi2,
                           const double *chi3, int nqp, double *mm) {
  static orc_tab t;
  if (orc_tab_init(&t, nqp)) return 1;
  const int64_t nc3 = (int64_t)ncells * nl;
  for (int cell = 0; cell < ncells; ++cell)
    for (int k = 0; k < nl; ++k) {
      const int64_t ik = (int64_t)cell * nl + k;
      double c[3][8], J[3][3], s[2][2] = {{0}};
      orc_cell_coords(&map[(size_t)cell * 8], k, chi1, chi2, chi3, c);
      for (int kz = 0; kz < nqp; ++kz)
        for (int ic = 0; ic < t.nh; ++ic) {
          const double dj = orc_jacobian(&t, c, ic, kz, J);
          const double b[2] = {orc_w2_value(&t, 4, ic, kz), orc_w2_value(&t, 5, ic, kz)};
          for (int df2 = 0; df2 < 2; ++df2)
            for (int df = df2; df < 2; ++df)
              s[df][df2] = s[df][df2] + t.wh[ic] * t.w[kz] * b[df] * b[df2] * dj;
        }
      for (int df2 = 0; df2 < 2; ++df2) {
        for (int df = df2; df < 2; ++df) mm[ik + nc3 * (df + 2 * (int64_t)df2)] = s[df][df2];
        for (int df = df2 - 1; df >= 0; --df) mm[ik + nc3 * (df + 2 * (int64_t)df2)] = s[df2][df];
      }
    }
  return 0;
}

/* Lowest-order W3 mass matrix: compute_mass_matrix_w3_code_r_double
 * (components/science/source/kernel/fem/sci_compute_mass_matrix_kernel_w3_mod.F90:270-356): the
 * one W3 basis function is 1, so the entry is the sum over (qp_v outer, qp_h inner) of
 * wh*wv*1*1*dj with rehabilitation, wh*wv*1*1/dj without.  mm has ncell_3d entries. */
int orc_mass_matrix_w3(int ncells, int nl, const int *map, const double *chi1, const double *chi2,
                       const double *chi3, int nqp, int rehabilitate, double *mm) {
  static orc_tab t;
  if (orc_tab_init(&t, nqp)) return 1;
  for (int cell = 0; cell < ncells; ++cell)
    for (int k = 0; k < nl; ++k) {
      double c[3][8], J[3][3], s = 0.0;
      orc_cell_coords(&map[(size_t)cell * 8], k, chi1, chi2, chi3, c);
      for (int kz = 0; kz < nqp; ++kz)
        for (int ic = 0; ic < t.nh; ++ic) {
          const double dj = orc_jacobian(&t, c, ic, kz, J);
          const double w = t.wh[ic] * t.w[kz] * 1.0 * 1.0;
          s = s + (rehabilitate ? w * dj : w / dj);
        }
      mm[(int64_t)cell * nl + k] = s;
    }
  return 0;
}

/* Weak divergence W2 -> W3: compute_div_operator_code
 * (components/science/source/kernel/fem/sci_compute_div_operator_kernel_mod.F90:120-150): entry
 * (1, df2) is the sum over (qp_v outer, qp_h inner) of wh*wv*1*div(b_df2) with rehabilitation,
 * that over dj without.  div in the reference layout (ncell_3d, 1, 6). */
int orc_div_operator(int ncells, int nl, const int *map, const double *chi1, const double *chi2,
                     const double *chi3, int nqp, int rehabilitate, double *div) {
  static orc_tab t;
  if (orc_tab_init(&t, nqp)) return 1;
  const int64_t nc3 = (int64_t)ncells * nl;
  for (int cell = 0; cell < ncells; ++cell)
    for (int k = 0; k < nl; ++k) {
      const int64_t ik = (int64_t)cell * nl + k;
      double c[3][8], J[3][3], s[6] = {0};
      orc_cell_coords(&map[(size_t)cell * 8], k, chi1, chi2, chi3, c);
      for (int kz = 0; kz < nqp; ++kz)
        for (int ic = 0; ic < t.nh; ++ic) {
          const double dj = orc_jacobian(&t, c, ic, kz, J);
          for (int d = 0; d < 6; ++d) {
            const double v = t.wh[ic] * t.w[kz] * 1.0 * orc_w2_div[d];
            s[d] = s[d] + (rehabilitate ? v : v / dj);
          }
        }
      for (int d = 0; d < 6; ++d) div[ik + nc3 * d] = s[d];
    }
  return 0;
}
