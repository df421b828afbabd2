/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, CPU-only restatement of the LFRic Core LMA-apply and Krylov-solve path, used
 * ONLY by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing in the product (lfric_core_amd/, include/) links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * lfric_core source tree).  Integer arrays use the reference's 1-based ids as values;
 * C indexing is 0-based.
 *
 * Parity pins (what checks each restatement against something other than itself):
 *   orc_cubedsphere_topology  netCDF UGRID fixtures C4/C8/C12/C16/C24/C32 (face_nodes, face_links)
 *   orc_local_vert_lids       invariant: lid == first encounter; nverts == 6n^2+2
 *   orc_cubedsphere_lonlat    node longitudes / latitudes of the reference's UGRID files C4, C12
 *                             (degrees, to 1e-10); every vertex on the unit sphere at the
 *                             equiangular grid's angles (a closed form per panel)
 *   orc_eta                   extrusion_mod_test.pf values (uniform, quadratic, geometric);
 *                             closed forms k/nl and (k/nl)^2
 *   orc_chi_w0                |chi| = scaled_radius + eta(k) (top - surface); chi / |chi| = the
 *                             vertex's unit vector (lon, lat)
 *   orc_mesh_connectivity     canned bi-periodic 3x3 mesh edge_on_cell / face_on_cell tables
 *   orc_dofmap                fs0k0 / fs1k0 / fs2k0 / fs3k0 dofmap fixtures; undf closed forms;
 *                             W0 order-1 entity sharing and column lengths
 *   orc_colour_cubedsphere    closed form + validator invariants + brute-force shared-vertex check
 *   orc_rng_*                 libm log/exp to a few ulp; N(0,1) / U(-1,1) moments
 *   orc_lma_apply             known-answer test of matrix_vector_mod_test.pf; assembled
 *                             sparse matrix (scipy) brute force; oracle/_ref (the reference
 *                             Fortran kernel compiled from source) bit-level comparison
 *   orc_cg / orc_gmres        numpy.linalg.solve on tiny SPD/nonsymmetric systems; scipy cg
 *                             solutions; Givens least-squares vs numpy.linalg.lstsq
 *   orc_colour_generic        greedy rule checked from vertex lists (independent neighbour sets)
 *   orc_colour_planar         valid 4-colouring; along W/E the lowest bit of colour-1 alternates
 *                             and along S/N the next one (the structure the transition table
 *                             gives), on the reference's planar UGRID meshes
 *   orc_transpose_apply,      known-answer tests of transpose/dg/dg_inc/mm_diagonal unit tests;
 *   orc_dg_apply,             bit-level comparison with the compiled reference kernels
 *   orc_dg_inc_apply,         (dg_inc also: accumulates onto lhs, y0 + D x)
 *   orc_mm_diagonal
 *   orc_enforce_bc            bit-level comparison with the compiled reference enforce_bc kernel;
 *   orc_boundary_dofs         W2: exactly the bottom and top dofs of the vertical-flux columns
 *                             (the domain's bottom and top surfaces) become zero, nothing else
 *                             changes; a flag on a side face fails that
 *   orc_gauss_rule            numpy.polynomial.legendre.leggauss mapped to [0, 1]; the 3-point
 *                             weights of get_unit_test_q3x3x3_quadrature_mod.f90
 *   orc_mass_matrix_w0        known answer of compute_mass_matrix_kernel_w0_mod_test.pf (a
 *                             6000 x 1000 x 2000 box); closed forms V (M1 x M1 x M1) for boxes
 *                             and det(J) M_ref for sheared parallelepipeds; sum of entries =
 *                             cell volume
 *   orc_w2_basis              basis and divergence tables of get_unit_test_q3x3x3_basis_mod.f90
 *   orc_mass_matrix_w2        known answer of compute_mass_matrix_kernel_w2_mod_test.pf; closed
 *                             form G(c_i, c_j) I_ij / det(A) with G = A^T A for sheared
 *                             parallelepipeds x = A xi + o
 *   orc_mass_matrix_w3        known answer of compute_mass_matrix_kernel_w3_mod_test.pf (volume);
 *                             equals the sum of the cell's W0 mass entries; independent of nqp
 *                             >= 2 on trilinear cells; 1/det(A) without rehabilitation
 *   orc_mass_matrix_wtheta    known answer of compute_mass_matrix_kernel_wtheta_mod_test.pf
 *                             (1/3, 1/6 of the volume); det(A) [[1/3,1/6],[1/6,1/3]] on sheared
 *                             parallelepipeds; entries add up to the W3 entry (the volume)
 *   orc_div_operator          known answer of compute_div_operator_kernel_mod_test.pf
 *                             (-1, 1, 1, -1, -1, 1); without rehabilitation = divergence sign x
 *                             the W3 entry without rehabilitation, bit for bit
 *   orc_cma_*                 known answers of the columnwise_op_{asm,app,appinv,mul,scaledadd,
 *   orc_tri_solve             asm_diag_hmht,asm_m2v_lumped_inv}_kernel_mod_test.pf and
 *                             sci_tri_solve_kernel_mod_test.pf; dense numpy restatements of the
 *                             banded products and solves; bit-level comparison with the
 *                             reference's CMA kernels compiled from source (r_solver = real32)
 *   orc_invert_local          known answers of invert_local_operator_kernel_mod_test.pf (1/5) and
 *                             matrix_invert_mod_test.pf (a a^-1 = I); numpy.linalg.inv on
 *                             nonsymmetric matrices; bit-level comparison with the reference's
 *                             matrix_invert_lu compiled from source
 *   orc_bicgstab              numpy.linalg.solve on a tiny nonsymmetric system
 *   orc_fgmres / orc_gcr      one cycle = minimal residual over the right-preconditioned Krylov
 *                             space (numpy lstsq on an explicit basis); GCR(m) == FGMRES(m+1)
 *   orc_jacobi                closed form x_k = x* + (I - rho P^-1 A)^k (x_0 - x*)
 *   orc_chebyshev             convergence to numpy.linalg.solve with exact eigenvalue bounds;
 *                             linearity in b
 *   orc_prec_only             x = b / diag exactly
 */
#ifndef LFRIC_ORACLE_H
#define LFRIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mesh / topology ---------------------------------------------------------------- */
int orc_cubedsphere_topology(int n, int *cell_next, int *verts_on_cell);
int orc_cubedsphere_lonlat(int n, double *lon, double *lat);
int orc_eta(int kind, int nl, double *eta);
int orc_chi_w0(int ncells, int nl, const int *dofmap, const int *vert_on_cell, const double *vlon,
               const double *vlat, int kind, double domain_top, double scaled_radius, double *chi1,
               double *chi2, double *chi3);
int orc_local_vert_lids(int ncells, const int *verts_on_cell_gid, int *vert_on_cell_lid);
int orc_mesh_connectivity(int ncells, const int *cell_next, const int *vert_on_cell,
                          int *face_on_cell, int *edge_on_cell, int *nfaces_2d,
                          int *nedges_2d);

/* function space ids (match the reference's fs_continuity enumeration by meaning only) */
enum { ORC_W0 = 0, ORC_W1 = 1, ORC_W2 = 2, ORC_W3 = 3 };
int orc_ndf(int space, int order);
int orc_dofmap(int space, int order, int ncells, int nlayers, int nverts_2d, int nedges_2d,
               const int *vert_on_cell, const int *face_on_cell, const int *edge_on_cell,
               int *dofmap, int *undf);

int orc_colour_cubedsphere(int n, const int *cell_next, int *colour_map);
int orc_colour_generic(int ncells, const int *cell_next, int *colour_map);
int orc_colour_planar(int ncells, const int *cell_next, int *colour_map);
int orc_colour_lookup(int ncells, int ncolours, const int *colour_map, int *npc, int *cic, int maxcc);
int orc_colour_validate(int ncells, const int *cell_next, const int *colour_map);

/* ---- counter-based synthetic inputs -------------------------------------------------- */
uint64_t orc_rng_bits(uint64_t seed, uint64_t stream, uint64_t ctr);
double orc_rng_uniform(uint64_t seed, uint64_t stream, uint64_t ctr);
double orc_rng_normal(uint64_t seed, uint64_t stream, uint64_t ctr);
double orc_log(double x);
double orc_exp(double x);
void orc_gen_spd(int ndf, int64_t ncell_3d, uint64_t seed, int scaled, double *matrix);
void orc_gen_spd_shift(int ndf, int64_t ncell_3d, uint64_t seed, int scaled, double shift,
                       double *matrix);
void orc_gen_uniform(int64_t n, uint64_t seed, uint64_t stream, double *v);
void orc_gen_op_uniform(int n1, int n2, int64_t ncell_3d, uint64_t seed, uint64_t stream, double lo,
                        double hi, const double *sign, double *matrix);

/* ---- LMA kernels ---------------------------------------------------------------------- */
void orc_lma_apply(int ncells, int nlayers, int ncolours, const int *ncells_per_colour,
                   const int *cells_in_colour, int maxcc, int ndf1, int undf1,
                   const int *map1, int ndf2, int undf2, const int *map2,
                   const double *matrix, const double *x, double *y);
void orc_lma_apply_r32(int ncells, int nlayers, int ncolours, const int *ncells_per_colour,
                       const int *cells_in_colour, int maxcc, int ndf1, int undf1,
                       const int *map1, int ndf2, int undf2, const int *map2,
                       const float *matrix, const float *x, float *y);
void orc_transpose_apply(int ncells, int nlayers, int ncolours, const int *ncells_per_colour,
                         const int *cells_in_colour, int maxcc, int ndf1, int undf1,
                         const int *map1, int ndf2, int undf2, const int *map2,
                         const double *matrix, const double *x, double *y);
void orc_dg_apply(int ncells, int nlayers, int ndf1, int undf1, const int *map1, int ndf2,
                  int undf2, const int *map2, const double *matrix, const double *x,
                  double *y);
void orc_dg_inc_apply(int ncells, int nlayers, int ndf1, int undf1, const int *map1, int ndf2,
                      int undf2, const int *map2, const double *matrix, const double *x,
                      double *y);
void orc_enforce_bc(int ncells, int nlayers, int ndf, const int *map, const int *bv, double *field);
int orc_boundary_dofs(int space, int order, int ndf, int *bv);
void orc_mm_diagonal(int ncells, int nlayers, int ndf, int undf, const int *map,
                     const double *matrix, double *diag);
/* per-cell dense inverse of a square operator (reference layout), FP64 and FP32 arithmetic */
#define ORC_MAX_NDF 32
int orc_invert_local(int64_t ncell3d, int ndf, const double *matrix, double *inv);
int orc_invert_local_r32(int64_t ncell3d, int ndf, const float *matrix, float *inv);
/* Gauss-Legendre points and weights on [0, 1]; lowest-order W0 mass matrix from Cartesian
 * vertex coordinates (chi1..3 W0 fields), nqp Gauss points per direction */
int orc_gauss_rule(int n, double *pts, double *wts);
int orc_mass_matrix_w0(int ncells, int nlayers, const int *map, const double *chi1, const double *chi2,
                       const double *chi3, int nqp, double *mm);
/* Lowest-order W2 basis tables, W2 and W3 mass matrices and the weak divergence W2 -> W3 from
 * the same coordinate fields (map is the W0 dofmap of the coordinate space) */
int orc_w2_basis(int nqp, double *basis, double *div);
int orc_mass_matrix_w2(int ncells, int nlayers, const int *map, const double *chi1, const double *chi2,
                       const double *chi3, int nqp, double *mm);
int orc_mass_matrix_w3(int ncells, int nlayers, const int *map, const double *chi1, const double *chi2,
                       const double *chi3, int nqp, int rehabilitate, double *mm);
int orc_mass_matrix_wtheta(int ncells, int nlayers, const int *map, const double *chi1, const double *chi2,
                           const double *chi3, int nqp, double *mm);
int orc_div_operator(int ncells, int nlayers, const int *map, const double *chi1, const double *chi2,
                     const double *chi3, int nqp, int rehabilitate, double *div);

/* ---- columnwise (CMA) operators and the vertical tridiagonal solve (orc_cma.c) ---------- *
 * p[7] = {alpha, beta, gamma_m, gamma_p, bandwidth, nrow, ncol}; matrices (bandwidth, nrow,
 * ncells) in Fortran order; _r32 variants are the r_solver = real32 instances. */
int orc_cma_params(int nl, int ni_to, int nf_to, int ni_from, int nf_from, int custom, int alpha, int beta,
                   int gamma_m, int gamma_p, int *p);
void orc_cma_dofmaps(int nl, int ndf, int ni, int nf, const int *map_cell1, int *banded, int *indirection);
#define ORC_CMA_DECLS(T, SFX)                                                                                  \
  void orc_cma_asm##SFX(int ncells, int nl, int ndf_to, int ndf_from, const T *lma, const int *p,              \
                        const int *banded_to, const int *banded_from, T *cma);                                  \
  void orc_cma_apply##SFX(int ncells, const int *p, const T *cma, const T *x, T *lhs, int ndf1, const int *map1, \
                          const int *ind_to, int ndf2, const int *map2, const int *ind_from);                   \
  void orc_cma_appinv##SFX(int ncells, const int *p, const T *cma, const T *x, T *lhs, int ndf, const int *map,  \
                           const int *ind);                                                                     \
  void orc_cma_mul##SFX(int ncells, const int *pa, const T *A, const int *pb, const T *B, const int *pc, T *C);   \
  void orc_cma_scaledadd##SFX(int ncells, const int *pa, const T *A, const int *pb, const T *B, const int *pc,   \
                              T *C, T alpha, T beta);                                                           \
  void orc_cma_asm_diag_hmht##SFX(int ncells, int nl, int ndf_w3, int ndf_w2h, const T *Dh, const T *M2h,      \
                                  const int *p, const int *banded, T *cma);                                     \
  void orc_cma_asm_m2v_lumped_inv##SFX(int ncells, int nl, int ndf, const T *lma, const int *p,                 \
                                       const int *banded, T *cma);                                              \
  void orc_tri_solve##SFX(int ncells, int nl, const int *map, const T *x, const T *t0, const T *tp, const T *tm, \
                          T *y);
ORC_CMA_DECLS(double, )
ORC_CMA_DECLS(float, _r32)
#undef ORC_CMA_DECLS

/* ---- Krylov solvers ------------------------------------------------------------------- */
typedef struct orc_problem {
  /* operator y = A x (LMA, W->W) */
  int ncells, nlayers, ncolours, maxcc, ndf, undf;
  const int *map, *ncells_per_colour, *cells_in_colour;
  const double *matrix;
  const double *diag; /* Jacobi diagonal, NULL for identity preconditioner */
  /* optional callbacks replacing the LMA operator / diagonal preconditioner */
  void (*apply_fn)(const void *ctx, const double *x, double *y);
  void (*prec_fn)(const void *ctx, const double *x, double *y);
  const void *ctx;
} orc_problem;

typedef struct orc_result {
  int iterations;
  int converged;
  double r_nrm_0;
  double r_nrm;
} orc_result;

int orc_cg(const orc_problem *p, const double *b, double *x, double r_tol, double a_tol,
           int max_iter, int monitor, orc_result *res, double *history);
int orc_gmres(const orc_problem *p, const double *b, double *x, int gcrk, double r_tol,
              double a_tol, int max_iter, int monitor, orc_result *res, double *history);
int orc_givens_lsq(int m, const double *h_in, const double *g_in, double *u);
int orc_bicgstab(const orc_problem *p, const double *b, double *x, double r_tol, double a_tol,
                 int max_iter, int monitor, orc_result *res, double *history);
int orc_fgmres(const orc_problem *p, const double *b, double *x, int gcrk, double r_tol,
               double a_tol, int max_iter, int monitor, orc_result *res, double *history);
int orc_gcr(const orc_problem *p, const double *b, double *x, int gcrk, double r_tol, double a_tol,
            int max_iter, int monitor, orc_result *res, double *history);
int orc_jacobi(const orc_problem *p, const double *b, double *x, double r_tol, double a_tol,
               int max_iter, int monitor, double rho_relax, orc_result *res, double *history);
int orc_chebyshev(const orc_problem *p, const double *b, double *x, int max_iter, double lmin,
                  double lmax, orc_result *res);
int orc_prec_only(const orc_problem *p, const double *b, double *x, orc_result *res);

#ifdef __cplusplus
}
#endif
#endif
