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

/* bicgstab_solve (sci_iterative_solver_mod.f90:645-779).  The preconditioner acts on the
 * search directions (y = P^-1 t, z = P^-1 s), so x is updated with preconditioned vectors.
 * history[iter] = ||r|| when monitoring.  Returns 0 converged / 1 not converged. */
int orc_bicgstab(const orc_problem *p, const double *b, double *x, double r_tol, double a_tol,
                 int max_iter, int monitor, orc_result *res, double *history) {
  const int n = p->undf;
  double *r = calloc((size_t)n, 8), *r0 = calloc((size_t)n, 8), *pp = calloc((size_t)n, 8),
         *v = calloc((size_t)n, 8), *t = calloc((size_t)n, 8), *s = calloc((size_t)n, 8),
         *y = calloc((size_t)n, 8), *z = calloc((size_t)n, 8);
  double alpha, beta, rho, omega, rho_old, err = 0.0, sc_err, tt, ts;
  int converged = 0, iter = 0;
  memcpy(r, b, sizeof(double) * (size_t)n);                 /* :670-671 */
  op_apply(p, x, v);                                        /* :673-675 */
  for (int i = 0; i < n; ++i) r[i] = r[i] + (-1.0) * v[i];
  memcpy(r0, r, sizeof(double) * (size_t)n);                /* :676-677 */
  sc_err = sqrt(dot(n, r, r));
  res->r_nrm_0 = sc_err;
  res->r_nrm = sc_err;
  if (history) history[0] = sc_err;
  if (sc_err < EPS_RDEF) {                                  /* :681-686 */
    res->iterations = 0; res->converged = 1;
    goto done;
  }
  if (a_tol > sc_err) sc_err = a_tol;                       /* :688 */
  alpha = 1.0; omega = 1.0; rho_old = 1.0;                  /* :694-696 */
  memset(v, 0, sizeof(double) * (size_t)n);
  memset(pp, 0, sizeof(double) * (size_t)n);
  for (iter = 1; iter <= max_iter; ++iter) {
    rho = dot(n, r, r0);                                    /* :708 */
    beta = (rho / rho_old) * (alpha / omega);
    for (int i = 0; i < n; ++i) t[i] = r[i] + (-beta * omega) * v[i];
    prec_apply(p, t, y);
    for (int i = 0; i < n; ++i) pp[i] = beta * pp[i] + y[i];  /* p%aypx(beta,y) */
    op_apply(p, pp, v);
    alpha = rho / dot(n, r0, v);                            /* :716 */
    for (int i = 0; i < n; ++i) s[i] = r[i] + (-alpha) * v[i];
    prec_apply(p, s, z);
    op_apply(p, z, t);
    tt = dot(n, t, t);                                      /* :722-724 */
    ts = dot(n, t, s);
    omega = ts / tt;
    for (int i = 0; i < n; ++i) x[i] = x[i] + omega * z[i];
    for (int i = 0; i < n; ++i) x[i] = x[i] + alpha * pp[i];
    for (int i = 0; i < n; ++i) r[i] = s[i] + (-omega) * t[i];
    rho_old = rho;
    if (monitor) {
      double rn = sqrt(dot(n, r, r));
      err = rn / sc_err;
      res->r_nrm = rn;
      if (history) history[iter] = rn;
      if (err < r_tol) {
        converged = 1;
        break;
      }
    }
  }
  res->iterations = converged ? iter : max_iter;
  res->converged = converged;
done:
  free(r); free(r0); free(pp); free(v); free(t); free(s); free(y); free(z);
  return res->converged ? 0 : 1;
}

/* fgmres_solve (sci_iterative_solver_mod.f90:1055-1237): flexible (right-preconditioned)
 * restarted GMRES.  The residual starts from b (the reference assumes a zero initial guess,
 * :1080), the correction dx is accumulated from the preconditioned basis Pv and added to x at
 * the end (:1224). */
int orc_fgmres(const orc_problem *p, const double *b, double *x, int gcrk, double r_tol,
               double a_tol, int max_iter, int monitor, orc_result *res, double *history) {
  const int n = p->undf, m = gcrk;
  double *dx = calloc((size_t)n, 8), *Ax = calloc((size_t)n, 8), *r = calloc((size_t)n, 8),
         *s = calloc((size_t)n, 8), *w = calloc((size_t)n, 8), *v = calloc((size_t)n * m, 8),
         *pv = calloc((size_t)n * m, 8);
  double *h = calloc((size_t)(m + 1) * m, 8), *g = calloc((size_t)m + 1, 8), *u = calloc((size_t)m, 8);
  double sc_err, err = 0.0, beta;
  int converged = 0, iter;
  memcpy(r, b, sizeof(double) * (size_t)n);
  sc_err = sqrt(dot(n, r, r));
  res->r_nrm_0 = sc_err;
  res->r_nrm = sc_err;
  if (history) history[0] = sc_err;
  if (sc_err < EPS_RDEF) {
    res->iterations = 0; res->converged = 1;
    goto done;
  }
  if (monitor && a_tol > sc_err) sc_err = a_tol;
  memcpy(s, r, sizeof(double) * (size_t)n);
  beta = sqrt(dot(n, s, s));
#define V(j) (&v[(size_t)(j) * n])
#define PV(j) (&pv[(size_t)(j) * n])
#define H(i, j) h[((j)-1) * (m + 1) + (i)-1]
  for (int i = 0; i < n; ++i) V(0)[i] = s[i] * (1.0 / beta);
  memset(g, 0, sizeof(double) * (size_t)(m + 1));
  g[0] = beta;
  for (iter = 1; iter <= max_iter; ++iter) {
    for (int iv = 1; iv <= m; ++iv) {
      prec_apply(p, V(iv - 1), PV(iv - 1));
      op_apply(p, PV(iv - 1), s);
      memcpy(w, s, sizeof(double) * (size_t)n);
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
      for (int i = 0; i < n; ++i) dx[i] = dx[i] + u[iv - 1] * PV(iv - 1)[i];
    op_apply(p, dx, Ax);
    for (int i = 0; i < n; ++i) r[i] = (-1.0) * Ax[i] + b[i];
    beta = sqrt(dot(n, r, r));
    res->r_nrm = beta;
    if (history) history[iter] = beta;
    if (monitor) {
      err = beta / sc_err;
      if (err < r_tol) {
        converged = 1;
        break;
      }
    }
    memcpy(s, r, sizeof(double) * (size_t)n);              /* :1203-1206 */
    for (int i = 0; i < n; ++i) V(0)[i] = 0.0;
    for (int i = 0; i < n; ++i) V(0)[i] = V(0)[i] + (1.0 / beta) * s[i];
    memset(g, 0, sizeof(double) * (size_t)(m + 1));
    g[0] = beta;
  }
#undef V
#undef PV
#undef H
  for (int i = 0; i < n; ++i) x[i] = x[i] + 1.0 * dx[i];
  res->iterations = converged ? iter : max_iter;
  res->converged = converged;
done:
  free(dx); free(Ax); free(r); free(s); free(w); free(v); free(pv); free(h); free(g); free(u);
  return res->converged ? 0 : 1;
}

/* gcr_solve (sci_iterative_solver_mod.f90:1288-1408): restarted generalised conjugate
 * residual with a right preconditioner.  dx starts as P^-1 b (:1310-1312), the scale is
 * ||b||, and x += dx at the end.  history[k] = ||res|| after the k-th basis vector. */
int orc_gcr(const orc_problem *p, const double *b, double *x, int gcrk, double r_tol, double a_tol,
            int max_iter, int monitor, orc_result *res, double *history) {
  const int n = p->undf, m = gcrk;
  double *dx = calloc((size_t)n, 8), *Ax = calloc((size_t)n, 8), *r = calloc((size_t)n, 8),
         *v = calloc((size_t)n * m, 8), *pv = calloc((size_t)n * m, 8);
  double alpha, beta, sc_err, err = 1.0;
  int converged = 0, iter, count = 0;
  prec_apply(p, b, dx);
  op_apply(p, dx, Ax);
  for (int i = 0; i < n; ++i) r[i] = b[i] + (-1.0) * Ax[i];
  sc_err = sqrt(dot(n, b, b));
  res->r_nrm_0 = sc_err;
  res->r_nrm = sqrt(dot(n, r, r));
  if (history) history[0] = res->r_nrm;
  if (sc_err < EPS_RDEF) {
    res->iterations = 0; res->converged = 1;
    goto done;
  }
  if (monitor && a_tol > sc_err) sc_err = a_tol;
#define V(j) (&v[(size_t)(j) * n])
#define PV(j) (&pv[(size_t)(j) * n])
  for (iter = 1; iter <= max_iter; ++iter) {
    for (int iv = 1; iv <= m; ++iv) {
      prec_apply(p, r, PV(iv - 1));
      op_apply(p, PV(iv - 1), V(iv - 1));
      for (int j = 1; j <= iv - 1; ++j) {
        alpha = dot(n, V(iv - 1), V(j - 1));
        for (int i = 0; i < n; ++i) V(iv - 1)[i] = V(iv - 1)[i] + (-alpha) * V(j - 1)[i];
        for (int i = 0; i < n; ++i) PV(iv - 1)[i] = PV(iv - 1)[i] + (-alpha) * PV(j - 1)[i];
      }
      alpha = sqrt(dot(n, V(iv - 1), V(iv - 1)));
      beta = 1.0 / alpha;
      for (int i = 0; i < n; ++i) V(iv - 1)[i] = V(iv - 1)[i] * beta;
      for (int i = 0; i < n; ++i) PV(iv - 1)[i] = PV(iv - 1)[i] * beta;
      alpha = dot(n, r, V(iv - 1));
      for (int i = 0; i < n; ++i) dx[i] = dx[i] + alpha * PV(iv - 1)[i];
      for (int i = 0; i < n; ++i) r[i] = r[i] + (-alpha) * V(iv - 1)[i];
      ++count;
      if (monitor) {
        double rn = sqrt(dot(n, r, r));
        err = rn / sc_err;
        res->r_nrm = rn;
        if (history) history[count] = rn;
        if (err < r_tol) break;
      }
    }
    if (monitor && err < r_tol) {
      converged = 1;
      break;
    }
  }
#undef V
#undef PV
  for (int i = 0; i < n; ++i) x[i] = x[i] + 1.0 * dx[i];
  res->iterations = converged ? iter : max_iter;
  res->converged = converged;
done:
  free(dx); free(Ax); free(r); free(v); free(pv);
  return res->converged ? 0 : 1;
}

/* jacobi_solve (sci_iterative_solver_mod.f90, jacobi_smod): x += -rho_relax P^-1 (A x - b),
 * convergence on ||A x - b|| relative to its first value. */
int orc_jacobi(const orc_problem *p, const double *b, double *x, double r_tol, double a_tol,
               int max_iter, int monitor, double rho_relax, orc_result *res, double *history) {
  const int n = p->undf;
  double *r = calloc((size_t)n, 8), *z = calloc((size_t)n, 8);
  double r_nrm = 0.0, r_nrm_0 = 0.0;
  const double c = -rho_relax;
  int converged = 0, iter;
  for (iter = 1; iter <= max_iter; ++iter) {
    op_apply(p, x, r);                                     /* r = A x */
    for (int i = 0; i < n; ++i) r[i] = r[i] + (-1.0) * b[i];
    if (monitor) {
      r_nrm = sqrt(dot(n, r, r));
      if (iter == 1) r_nrm_0 = r_nrm;
      if (history) history[iter - 1] = r_nrm;
      if (r_nrm / r_nrm_0 <= r_tol || r_nrm <= a_tol) {
        converged = 1;
        break;
      }
    }
    prec_apply(p, r, z);
    for (int i = 0; i < n; ++i) x[i] = x[i] + c * z[i];
  }
  res->r_nrm_0 = r_nrm_0;
  res->r_nrm = r_nrm;
  res->iterations = converged ? iter : max_iter;
  res->converged = converged;
  free(r); free(z);
  return converged ? 0 : 1;
}

/* chebyshev_solve (sci_iterative_solver_mod.f90, chebyshev_smod): max_iter steps of the
 * preconditioned Chebyshev recurrence for eigenvalue bounds [lmin, lmax], from x = 0.  The
 * returned r_nrm is ||A x - b|| and r_nrm_0 = max(1, ||b||) (the reference's scale). */
int orc_chebyshev(const orc_problem *p, const double *b, double *x, int max_iter, double lmin,
                  double lmax, orc_result *res) {
  const int n = p->undf;
  double *xp = calloc((size_t)n, 8), *xo = calloc((size_t)n, 8), *z = calloc((size_t)n, 8),
         *r = calloc((size_t)n, 8);
  const double a1 = 2.0 / (lmax - lmin), a2 = (lmax + lmin) / (lmax - lmin), a_over_b = a1 / a2;
  double w = 1.0;
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  res->r_nrm_0 = fmax(1.0, sqrt(dot(n, b, b)));
  for (int iter = 1; iter <= max_iter; ++iter) {
    op_apply(p, xo, z);
    for (int i = 0; i < n; ++i) r[i] = 1.0 * b[i] + (-1.0) * z[i];  /* r%axpby(1,b,-1,z) */
    prec_apply(p, r, z);
    w = 1.0 / (1.0 - w / (4.0 * a2 * a2));
    const double wab = w * a_over_b;
    for (int i = 0; i < n; ++i) x[i] = wab * z[i] + w * xo[i];      /* x%axpby(wab,z,w,xo) */
    for (int i = 0; i < n; ++i) x[i] = x[i] + (1.0 - w) * xp[i];
    if (iter < max_iter) {
      memcpy(xp, xo, sizeof(double) * (size_t)n);
      memcpy(xo, x, sizeof(double) * (size_t)n);
    }
  }
  op_apply(p, x, z);
  for (int i = 0; i < n; ++i) z[i] = z[i] + (-1.0) * b[i];
  res->r_nrm = sqrt(dot(n, z, z));
  res->iterations = max_iter;
  res->converged = 1;
  free(xp); free(xo); free(z); free(r);
  return 0;
}

/* precondition_only_solve (sci_iterative_solver_mod.f90, precondition_only_smod): x = P^-1 b;
 * r_nrm_0 = ||b||, r_nrm = ||b - A x||. */
int orc_prec_only(const orc_problem *p, const double *b, double *x, orc_result *res) {
  const int n = p->undf;
  double *Ax = calloc((size_t)n, 8), *r = calloc((size_t)n, 8);
  prec_apply(p, b, x);
  memcpy(r, b, sizeof(double) * (size_t)n);
  res->r_nrm_0 = sqrt(dot(n, r, r));
  op_apply(p, x, Ax);
  for (int i = 0; i < n; ++i) r[i] = r[i] + (-1.0) * Ax[i];
  res->r_nrm = sqrt(dot(n, r, r));
  res->iterations = 1;
  res->converged = 1;
  free(Ax); free(r);
  return 0;
}
