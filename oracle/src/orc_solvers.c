/*
 * ORACLE — test infrastructure only (see oracle.h).
 *
 * FP64 preconditioned CG and left-preconditioned restarted GMRES, restated in the operation
 * order of components/science/source/solver/sci_iterative_solver_mod.f90, over an LMA
 * operator (apply = setval_c(y,0) + coloured matrix_vector_code) and a diagonal (Jacobi)
 * preconditioner z = r / diag (components/science/source/algorithm/solver/
 * sci_diagonal_preconditioner_alg_mod.x90:71-106, X_divideby_Y).  Vector operations follow
 * the field_vector_type built-ins (sci_field_vector_mod.x90:132-457); inner products are
 * summed sequentially in dof order (one partition: the global sum is the identity).
 */
#include "oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* EPS of constants_mod (infrastructure/source/utilities/constants_mod.F90) is
 * epsilon(1.0_r_def) for r_def = real64 */
#define EPS_RDEF DBL_EPSILON

static void op_apply(const orc_problem *p, const double *x, double *y) {
  if (p->apply_fn) {
    p->apply_fn(p->ctx, x, y);
    return;
  }
  orc_lma_apply(p->ncells, p->nlayers, p->ncolours, p->ncells_per_colour, p->cells_in_colour,
                p->maxcc, p->ndf, p->undf, p->map, p->ndf, p->undf, p->map, p->matrix, x, y);
}
static void prec_apply(const orc_problem *p, const double *x, double *y) {
  if (p->prec_fn)
    p->prec_fn(p->ctx, x, y);
  else if (p->diag)
    for (int i = 0; i < p->undf; ++i) y[i] = x[i] / p->diag[i];
  else
    memcpy(y, x, sizeof(double) * (size_t)p->undf);
}
static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + a[i] * b[i];
  return s;
}

/* cg_solve (sci_iterative_solver_mod.f90:495-592).  history (optional) receives ||r||
 * per iteration (index 0 = initial).  Returns 0 converged / 1 not converged. */
int orc_cg(const orc_problem *p, const double *b, double *x, double r_tol, double a_tol,
           int max_iter, int monitor, orc_result *res, double *history) {
  const int n = p->undf;
  double *r = calloc((size_t)n, 8), *pp = calloc((size_t)n, 8), *z = calloc((size_t)n, 8);
  double alpha, beta, r_nrm = 0.0, r_nrm_0, rz, rz_new;
  int converged = 0, iter;
  op_apply(p, x, r);                                       /* r = A.x      :521 */
  for (int i = 0; i < n; ++i) r[i] = -1.0 * r[i];          /* r%scale(-1)  :522 */
  for (int i = 0; i < n; ++i) r[i] = r[i] + 1.0 * b[i];    /* r%axpy(1,b)  :523 */
  r_nrm_0 = sqrt(dot(n, r, r));                            /* :524 */
  res->r_nrm_0 = r_nrm_0;
  if (history) history[0] = r_nrm_0;
  if (r_nrm_0 < EPS_RDEF) {                                /* :527-532 */
    res->iterations = 0; res->converged = 1; res->r_nrm = r_nrm_0;
    free(r); free(pp); free(z);
    return 0;
  }
  prec_apply(p, r, z);                                     /* :541 */
  rz = dot(n, r, z);
  memcpy(pp, z, sizeof(double) * (size_t)n);
  for (iter = 1; iter <= max_iter; ++iter) {               /* :551-575 */
    op_apply(p, pp, z);                                    /* z = A.p */
    alpha = rz / dot(n, pp, z);
    for (int i = 0; i < n; ++i) x[i] = x[i] + alpha * pp[i];
    for (int i = 0; i < n; ++i) r[i] = r[i] + (-alpha) * z[i];
    if (monitor) {
      r_nrm = sqrt(dot(n, r, r));
      if (history) history[iter] = r_nrm;
      if (r_nrm / r_nrm_0 <= r_tol || r_nrm <= a_tol) {
        converged = 1;
        break;
      }
    }
    prec_apply(p, r, z);
    rz_new = dot(n, r, z);
    beta = rz_new / rz;
    for (int i = 0; i < n; ++i) pp[i] = beta * pp[i] + z[i]; /* p%aypx(beta,z) */
    rz = rz_new;
  }
  res->iterations = converged ? iter : max_iter;
  res->converged = converged;
  res->r_nrm = r_nrm;
  free(r); free(pp); free(z);
  return converged ? 0 : 1;
}

/* Givens rotations on the (m+1) x m Hessenberg h (column-major, leading dim m+1) and g,
 * then back substitution for u (sci_iterative_solver_mod.f90:929-953). */
int orc_givens_lsq(int m, const double *h_in, const double *g_in, double *u) {
  const int ld = m + 1;
  double *h = malloc(sizeof(double) * (size_t)ld * m), *g = malloc(sizeof(double) * (size_t)ld);
  memcpy(h, h_in, sizeof(double) * (size_t)ld * m);
  memcpy(g, g_in, sizeof(double) * (size_t)ld);
#define H(i, j) h[((j)-1) * ld + (i)-1]
  for (int iv = 1; iv <= m; ++iv) {
    double nrm = sqrt(H(iv, iv) * H(iv, iv) + H(iv + 1, iv) * H(iv + 1, iv));
    double si = H(iv + 1, iv) / nrm, ci = H(iv, iv) / nrm;
    double pq = ci * g[iv - 1] + si * g[iv];
    double q = -si * g[iv - 1] + ci * g[iv];
    g[iv - 1] = pq;
    g[iv] = q;
    for (int j = iv; j <= m; ++j) {
      double h1 = ci * H(iv, j) + si * H(iv + 1, j);
      double h2 = -si * H(iv, j) + ci * H(iv + 1, j);
      H(iv, j) = h1;
      H(iv + 1, j) = h2;
    }
  }
  u[m - 1] = g[m - 1] / H(m, m);
  for (int iv = m - 1; iv >= 1; --iv) {
    u[iv - 1] = g[iv - 1];
    for (int j = iv + 1; j <= m; ++j) u[iv - 1] = u[iv - 1] - H(iv, j) * u[j - 1];
    u[iv - 1] = u[iv - 1] / H(iv, iv);
  }
#undef H
  free(h); free(g);
  return 0;
}

/* gmres_solve (sci_iterative_solver_mod.f90:830-1003): left preconditioned, modified
 * Gram-Schmidt Arnoldi, full gcrk-step cycles, convergence on the true residual after each
 * cycle.  history (optional) receives ||b - A x|| after each cycle (index 0 = initial). */
int orc_gmres(const orc_problem *p, const double *b, double *x, int gcrk, double r_tol,
              double a_tol, int max_iter, int monitor, orc_result *res, double *history) {
  const int n = p->undf, m = gcrk;
  double *Ax = calloc((size_t)n, 8), *r = calloc((size_t)n, 8), *s = calloc((size_t)n, 8),
         *w = calloc((size_t)n, 8), *v = calloc((size_t)n * m, 8);
  double *h = calloc((size_t)(m + 1) * m, 8), *g = calloc((size_t)m + 1, 8), *u = calloc((size_t)m, 8);
  double sc_err, err = 0.0;
  int converged = 0, iter;
  op_apply(p, x, Ax);
  for (int i = 0; i < n; ++i) r[i] = b[i] + (-1.0) * Ax[i];
  sc_err = sqrt(dot(n, r, r));
  res->r_nrm_0 = sc_err;
  if (history) history[0] = sc_err;
  if (sc_err < EPS_RDEF) {
    res->iterations = 0; res->converged = 1; res->r_nrm = sc_err;
    goto done;
  }
  if (monitor && a_tol > sc_err) sc_err = a_tol;
#define V(j) (&v[(size_t)(j) * n])
#define H(i, j) h[((j)-1) * (m + 1) + (i)-1]
  for (iter = 1; iter <= max_iter; ++iter) {
    prec_apply(p, r, s);
    double beta = sqrt(dot(n, s, s));
    for (int i = 0; i < n; ++i) V(0)[i] = s[i] * (1.0 / beta);
    memset(g, 0, sizeof(double) * (size_t)(m + 1));
    g[0] = beta;
    for (int iv = 1; iv <= m; ++iv) {
      memcpy(w, V(iv - 1), sizeof(double) * (size_t)n);
      op_apply(p, w, s);
      prec_apply(p, s, w);
      for (int j = 1; j <= iv; ++j) {
        H(j, iv) = dot(n, V(j - 1), w);
        for (int i = 0; i < n; ++i) w[i] = w[i] + (-H(j, iv)) * V(j - 1)[i];
      }
      H(iv + 1, iv) = sqrt(dot(n, w, w));
      if (iv < m)
        for (int i = 0; i < n; ++i) V(iv)[i] = w[i] * (1.0 / H(iv + 1, iv));
    }
    orc_givens_lsq(m, h, g, u);
    for (int iv = 1; iv <= m; ++iv)
      for (int i = 0; i < n; ++i) x[i] = x[i] + u[iv - 1] * V(iv - 1)[i];
    op_apply(p, x, Ax);
    for (int i = 0; i < n; ++i) r[i] = (-1.0) * Ax[i] + b[i]; /* res%aypx(-1,b) */
    if (monitor) {
      double rn = sqrt(dot(n, r, r));
      err = rn / sc_err;
      if (history) history[iter] = rn;
      res->r_nrm = rn;
      if (err < r_tol) {
        converged = 1;
        break;
      }
    }
  }
#undef V
#undef H
  res->iterations = converged ? iter : max_iter;
  res->converged = converged;
done:
  free(Ax); free(r); free(s); free(w); free(v); free(h); free(g); free(u);
  return converged ? 0 : 1;
}
