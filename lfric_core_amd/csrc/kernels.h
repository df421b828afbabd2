// Kernel launchers of liblfric_hip (host side declarations).
#pragma once
#include "internal.h"

namespace lfric {

// column roles of a function space's cell dofs, passed by value to kernels
struct Roles {
  uint32_t top_mask;
  int8_t partner[32];
};
Roles roles_of(const FunctionSpace &fs);

// device solver state (FP64 scalars stay on the GPU between kernels)
struct SolverState {
  double rz, alpha, beta, r_nrm0, r_nrm, r_tol, a_tol, pz;
  double rzb[2];      // rz of even / odd iterations (read and written by different kernels' blocks)
  double red0, red1;  // last reduction results
  int iter, max_iter, done, status, monitor, converged;
  unsigned int counter;
  int pad;
};

constexpr int kRedBlocks = 2048;  // fixed grid of the reduction kernels (deterministic sums)

// generators
int launch_gen_spd(lfric_op *op, unsigned long long seed, int scaled, double shift);
int launch_gen_uniform(lfric_op *op, unsigned long long seed, unsigned long long stream,
                       double lo, double hi, const double *col_sign_host);
int launch_vec_uniform(lfric_vec *v, unsigned long long seed, unsigned long long stream);
// layout changes between the reference local_stencil and the device plane layout
// reference-layout transfer; ref_precision LFRIC_R64 (double) or LFRIC_R32 (float) names the
// type of d_ref, converted to or from the operator's own storage precision
int launch_op_from_reference(lfric_op *op, const void *d_ref, int ref_precision);
int launch_op_to_reference(const lfric_op *op, void *d_ref, int ref_precision);
int launch_op_convert(const lfric_op *src, lfric_op *dst);

// operator kernels
int launch_lma_apply(const lfric_op *op, const void *x, void *y, const int *d_skip);
int launch_lma_apply_gather(const lfric_op *op, const void *x, void *y);
// gather apply that also leaves one x.y partial sum per block in `partials` (square
// operators); stops at once when *skip is set
int launch_lma_apply_gather_dot(const lfric_op *op, const void *x, void *y, double *partials, const int *skip);
int launch_gather_blocks(const lfric_op *op, int *nblocks);
// blocks of the gather apply's grid, once the space's column lists exist, and the number of
// p.(A p) partials its CG form leaves
int gather_blocks(const lfric_op *op);
int gather_dot_blocks(const lfric_op *op);
int launch_build_rows(lfric_op *op);
// the gather apply with `mark` recorded between the two phases of the symmetric-packed form
// (two_phase tells whether it was)
int launch_lma_apply_gather_marked(const lfric_op *op, const void *x, void *y, hipEvent_t mark, bool *two_phase);
// the apply form the solvers use (lfric_set_apply_form)
int launch_lma_apply_solver(const lfric_op *op, const void *x, void *y, const int *skip);
int launch_lma_apply_mrhs(const lfric_op *op, const void *x, void *y, int nrhs, int variant);
int launch_op_invert_local(const lfric_op *A, lfric_op *Ainv);
int launch_build_sym(lfric_op *op, int levels_per_thread);
// free the operator's symmetric-packed copy and its buffers (if any)
void drop_sym(lfric_op *op);
// the operator's symmetric-packed copy is the order-1 W0 kind (sym27_apply.h)
bool sym27_ready(const lfric_op *op);
int launch_assemble_mass_w0(lfric_op *op, const double *chi1, const double *chi2, const double *chi3, int nqp);
// tridiagonal solve of every column of a square columnwise operator (columnwise_op_appinv)
int launch_cma_appinv(const lfric_cma *A, const void *x, void *y);
// kind 0: W3 mass, 1: W2 mass, 2: weak divergence W2 -> W3; chi on the W0 space chi_fs
int launch_assemble_fem(lfric_op *op, int kind, const FunctionSpace &chi_fs, const double *chi1, const double *chi2,
                        const double *chi3, int nqp, int rehab);
// whether the MFMA multi-RHS apply (LFRIC_MRHS_MFMA) handles this operator
bool mfma_apply_supported(const lfric_op *op);
// colour-ordered form of a batch problem's apply: 1 matrix cores (lowest-order W0), 2 one unit
// per thread with register-staged x (order-1 W0), 0 none (gather form)
int batch_colour_form(const lfric_op *op);
int launch_lma_transpose(const lfric_op *op, const void *x, void *y);
// the Helmholtz composite in two kernels (lowest-order W3 <- W2), equal to the chain of
// transpose, division, divergence, M3 apply and axpy bit for bit; helmholtz_fusable says
// whether it applies to the operator
bool helmholtz_fusable(const lfric_linop *L);
int launch_helmholtz_fused(const lfric_linop *L, const void *x, void *y);
int launch_dg_apply(const lfric_op *op, const void *x, void *y);
int launch_dg_inc_apply(const lfric_op *op, const void *x, void *y);
// enforce_bc on a vector of space f (no-op when the space flags no boundary dofs)
int launch_enforce_bc(const FunctionSpace &f, int prec, void *y, const int *skip);
int launch_mm_diagonal(const lfric_op *op, void *diag);

// vector kernels (precision of the vectors, n elements)
int launch_set(int prec, void *v, long long n, double s);
int launch_axpy(int prec, void *y, const void *x, long long n, double a);
int launch_aypx(int prec, void *y, const void *x, long long n, double a);
int launch_axpby(int prec, void *z, double a, const void *x, double b, const void *y, long long n);
int launch_scale(int prec, void *v, long long n, double a);
int launch_copy(int prec, void *dst, const void *src, long long n);
int launch_divide(int prec, void *z, const void *x, const void *d, long long n);
int launch_convert(int pdst, void *dst, int psrc, const void *src, long long n, double scale);
int launch_axpy_mixed(double *y, double a, int prec_x, const void *x, long long n);
int launch_dot(int prec, const void *x, const void *y, long long n, double *d_out);

// CG building blocks on a SolverState
int launch_cg_residual(int prec, void *r, const void *b, long long n);  // r = b - r
int launch_cg_init_norm(int prec, const void *r, long long n, SolverState *st);
int launch_cg_init_prec(int prec, const void *r, void *z, void *p, const void *diag, long long n,
                        SolverState *st);
// p.(A p) partials (kRedBlocks) into `partials`
int launch_cg_dot_pz(int prec, const void *p, const void *z, long long n, SolverState *st, double *partials);
// x += alpha p, r -= alpha A p, z = P^-1 r and the (r.r, r.z) partials.  With pz_partials the
// kernel forms alpha from those npz partials itself; otherwise it reads alpha from st.  par:
// the parity of the iteration (its rz slot)
int launch_cg_update(int prec, void *x, void *r, void *z, const void *p, const void *diag, long long n,
                     SolverState *st, const double *pz_partials, int npz, int par);
// convergence test, beta and p = z + beta p from the update's partials
int launch_cg_aypx(int prec, void *p, const void *z, long long n, SolverState *st, int par);
int launch_cg_finish_pz(const double *partials, int nblocks, SolverState *st, int par);
// partials the folded update adds up itself, at most
constexpr int kFoldMaxPartials = 4 * kRedBlocks;

// CG on an LMA linear operator, with the mass-matrix operator's boundary conditions when it
// applies them (api.hip)
int cg_solve_linop(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                   const lfric_solver_opts *opts, lfric_solver_result *res);

// abstract operators and preconditioners (krylov.hip)
int linop_apply(const lfric_linop *L, const void *x, void *y);
int prec_apply_any(const lfric_prec *P, int prec, const void *x, void *y, long long n);

}  // namespace lfric
