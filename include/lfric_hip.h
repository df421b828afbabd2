/*
 * lfric_hip.h — C ABI of the MI355X-native LMA operator-apply and Krylov-solve library
 * (liblfric_hip.so, built from lfric_core_amd/csrc).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  Field and operator
 * data live in GPU memory owned by the handles; host data crosses only in the *_upload /
 * *_download calls (the import_field / export_field points of the reference's
 * field_vector_type, components/science/source/algorithm/solver/sci_field_vector_mod.x90:294-334).
 *
 * Every entry point names the reference interface it replaces.  Reference paths are relative
 * to the lfric_core source tree.
 *
 * Status codes (replace log_event(..., LOG_LEVEL_ERROR) aborts, infrastructure/source/
 * utilities/log_mod.F90:423-471, and the solver's converged / not-converged / NaN exits,
 * components/science/source/solver/sci_iterative_solver_mod.f90:577-590, 980-988).
 */
#ifndef LFRIC_HIP_H
#define LFRIC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define LFRIC_OK 0
#define LFRIC_NOT_CONVERGED 1 /* max_iter reached without meeting r_tol / a_tol */
#define LFRIC_NAN 2           /* NaN in a residual norm */
#define LFRIC_BREAKDOWN 3     /* zero / negative curvature or zero Krylov norm */
#define LFRIC_E_ARG (-1)      /* invalid argument or shape mismatch */
#define LFRIC_E_HIP (-2)      /* HIP runtime error */
#define LFRIC_E_NOMEM (-3)    /* device allocation failed */
#define LFRIC_E_STATE (-4)    /* library not initialised / handle misuse */
#define LFRIC_E_UNSUPPORTED (-5)

/* storage precisions: r_def / r_solver kinds of infrastructure/source/utilities/
 * constants_mod.F90:55-61 (r_solver = real32 by default, infrastructure/build/lfric.mk:126) */
#define LFRIC_R64 0
#define LFRIC_R32 1
#define LFRIC_BF16 2
#define LFRIC_F16 3

/* function spaces of infrastructure/source/function_space/fs_continuity_mod (by meaning) */
#define LFRIC_W0 0
#define LFRIC_W2 2
#define LFRIC_W3 3

typedef struct lfric_fs lfric_fs;     /* mesh + function space: dofmap, colouring (device) */
typedef struct lfric_op lfric_op;     /* LMA operator local_stencil (device) */
typedef struct lfric_vec lfric_vec;   /* field vector data (device) */
typedef struct lfric_prec lfric_prec; /* preconditioner */
typedef struct lfric_linop lfric_linop; /* abstract linear operator (LMA or composite) */
typedef struct lfric_cma lfric_cma;   /* columnwise (CMA) banded operator (device) */

/* ---- library ------------------------------------------------------------------------ */
/* Select the device and create the library stream (default priority). */
int lfric_init(int device);
/* Last error message of the calling thread ("" when none). */
const char *lfric_last_error(void);
/* Wait for all work queued by the library. */
int lfric_synchronize(void);
/* Launch this host thread's library calls on `stream` (a hipStream_t the caller created, at
 * default priority; NULL: the library's own stream), so that one host thread per stream drives
 * its own work (the threading of a serving loop).  Calls that wait for results wait on that
 * stream.  Applies and vector operations on different streams may overlap; the solvers share
 * the library's device scratch and state, so their calls must not overlap one another.  No
 * reference counterpart. */
int lfric_set_stream(void *stream);
/* Time work on the library stream with HIP events: start, then stop returns milliseconds. */
int lfric_timer_start(void);
int lfric_timer_stop(double *ms);
/* Bytes of device memory held by live handles. */
long long lfric_device_bytes(void);
/* LMA apply used inside the solvers: 0 the colour-ordered scatter (4 launches per apply),
 * 1 the gather form (one launch; in CG it also forms p.(A p)).  Default 1. */
int lfric_set_apply_form(int form);
/* Free and total device memory (hipMemGetInfo) of the library's device. */
int lfric_device_memory(long long *free_bytes, long long *total_bytes);

/* ---- function spaces (mesh_type + function_space_type inputs of the hot path) ---------- */
/* Wrap a caller-held dofmap and colouring.  Replaces the reads of
 * function_space_type%get_whole_dofmap (infrastructure/source/function_space/
 * function_space_mod.F90:748-757) and mesh_type%get_colour_map /
 * get_last_halo_cell_per_colour (infrastructure/source/mesh/mesh_mod.F90:2202-2245) made by
 * the PSy layer.  dofmap is (ndf, ncells) column-major with 1-based dof ids at the column
 * base; cells_in_colour is (maxcc, ncolours) column-major with 1-based cell ids. */
int lfric_fs_create(int ncells, int nlayers, int ndf, int undf, const int *dofmap,
                    int ncolours, const int *ncells_per_colour, const int *cells_in_colour,
                    int maxcc, lfric_fs **fs);
/* Build the serial-partition extruded cubed-sphere C<edge_cells> function space of the
 * reference's mesh generator and dofmap setup (mesh_tools/source/support/gencube_ps_mod.F90:
 * 336-720, infrastructure/source/function_space/
 * function_space_constructor_helper_functions_mod.F90:1715-2300) and its 4-colouring
 * (infrastructure/source/mesh/mesh_colouring_mod.F90:151-308). */
int lfric_fs_create_cubedsphere(int edge_cells, int nlayers, int space, int order,
                                lfric_fs **fs);
/* Same mesh as `like` (same cells, layers and colouring), another space/order. */
int lfric_fs_create_sibling(const lfric_fs *like, int space, int order, lfric_fs **fs);
/* The global mesh of a UGRID 2D file (infrastructure/source/io/ugrid_2d_mod.F90 reads the same
 * variables; netCDF classic or 64-bit-offset format): the mesh topology variable mesh_name (NULL:
 * the file's only one), its face_node_connectivity and face_face_connectivity as 1-based
 * [face][4] arrays (0 = no neighbour) and its npanels attribute (0 if absent).  Host only (no
 * device needed); call with null arrays to get nfaces first. */
int lfric_ugrid_read(const char *path, const char *mesh_name, int *nfaces, int *npanels, int *face_nodes,
                     int *face_links);
/* A function space on the extruded mesh of a UGRID file, built as for a generated mesh (local
 * vertex ids, 3D connectivity, dofmap_setup) and coloured as mesh_colouring_mod.F90:49-128 does:
 * the cubed-sphere closed form for six panels, the planar breadth-first form for one panel of
 * quadrilaterals, the generic greedy form otherwise. */
int lfric_fs_create_ugrid(const char *path, const char *mesh_name, int nlayers, int space, int order,
                          lfric_fs **fs);
int lfric_fs_info(const lfric_fs *fs, int *ncells, int *nlayers, int *ndf, int *undf,
                  int *ncolours);
int lfric_fs_get_dofmap(const lfric_fs *fs, int *dofmap);
int lfric_fs_get_colouring(const lfric_fs *fs, int *ncells_per_colour, int *cells_in_colour);
/* get_boundary_dofs (infrastructure/source/function_space/function_space_mod.F90:791-801):
 * boundary_dofs(ndf, 2) in Fortran order, 0 for a dof on the bottom (column 1) or top (column 2)
 * boundary, else 1.  Generated spaces carry the reference's flags
 * (function_space_constructor_helper_functions_mod.F90:678-680, 1048-1049: lowest-order W2
 * flags faces B and T; W0 and W3 none); the setter gives a caller-built space its own. */
int lfric_fs_get_boundary_dofs(const lfric_fs *fs, int *boundary_dofs);
int lfric_fs_set_boundary_dofs(lfric_fs *fs, const int *boundary_dofs);
void lfric_fs_destroy(lfric_fs *fs);

/* extrusions of infrastructure/source/mesh/extrusion_mod.f90: uniform k/nl (:292-305),
 * quadratic (k/nl)^2 (:341-354), geometric with stretching 1.03 (:390-409, :618-629) */
#define LFRIC_EXTRUSION_UNIFORM 0
#define LFRIC_EXTRUSION_QUADRATIC 1
#define LFRIC_EXTRUSION_GEOMETRIC 2
/* eta(0:nlayers) of an extrusion (the extrude methods of extrusion_mod.f90). */
int lfric_extrusion_eta(int kind, int nlayers, double *eta);
/* The Cartesian coordinate fields chi1..3 (FP64 vectors on fs) of a generated lowest-order W0
 * cubed-sphere space: vertex longitude/latitude of the mesh generator
 * (mesh_tools/source/support/gencube_ps_mod.F90:887-1113), levels from the extrusion over a
 * zero surface height to domain_top (mesh_constructor_helper_functions_mod.F90:120-252,
 * 450-477), projected onto spheres of radius scaled_radius + height as assign_coordinate_field
 * does for Cartesian coordinates (components/driver/source/driver_coordinates_mod.F90:59-290,
 * 357-454).  Computed on the host, uploaded once. */
int lfric_fs_coordinates(const lfric_fs *fs, int extrusion, double domain_top, double scaled_radius,
                         lfric_vec *chi1, lfric_vec *chi2, lfric_vec *chi3);

/* ---- field vectors (vector_mod abstract_vector_type, infrastructure/source/field/
 *      vector_mod.F90:17-59; field_vector_type / r_solver_field_vector_type) ----------- */
int lfric_vec_create(const lfric_fs *fs, int precision, int nrhs, lfric_vec **v);
int lfric_vec_upload(lfric_vec *v, const double *host);    /* import_field */
int lfric_vec_download(const lfric_vec *v, double *host);  /* export_field */
int lfric_vec_upload_r32(lfric_vec *v, const float *host);
int lfric_vec_download_r32(const lfric_vec *v, float *host);
int lfric_vec_size(const lfric_vec *v, int *undf, int *nrhs); /* vector_size */
void lfric_vec_destroy(lfric_vec *v);
int lfric_vec_set_scalar(lfric_vec *v, double scalar);                /* setval_c */
int lfric_vec_axpy(lfric_vec *y, double alpha, const lfric_vec *x);   /* y = y + a x */
int lfric_vec_aypx(lfric_vec *y, double alpha, const lfric_vec *x);   /* y = a y + x */
int lfric_vec_axpby(lfric_vec *z, double a, const lfric_vec *x, double b,
                    const lfric_vec *y);                              /* z = a x + b y */
int lfric_vec_scale(lfric_vec *v, double alpha);                      /* v = a v */
int lfric_vec_copy(lfric_vec *dst, const lfric_vec *src);             /* setval_X */
int lfric_vec_divide(lfric_vec *z, const lfric_vec *x, const lfric_vec *d); /* X_divideby_Y */
/* FP64-accumulated inner products (components/science/source/psy/sci_psykal_light_mod.f90:
 * 72-214; single partition so the global sum is the identity). */
int lfric_vec_dot(const lfric_vec *x, const lfric_vec *y, double *out);
int lfric_vec_norm(const lfric_vec *x, double *out);
/* Fill with U(-1,1) from the counter-based generator (seed, stream), dof index counter. */
int lfric_vec_fill_uniform(lfric_vec *v, unsigned long long seed, unsigned long long stream);

/* ---- LMA operators (operator_type, infrastructure/source/operator/operator_mod.t90:31-137) */
int lfric_op_create(const lfric_fs *to, const lfric_fs *from, int precision, lfric_op **op);
/* local_stencil(ncell_3d, ndf_to, ndf_from), column-major, FP64. */
int lfric_op_upload(lfric_op *op, const double *local_stencil);
int lfric_op_download(const lfric_op *op, double *local_stencil);
/* The same in real32: r_solver_operator_type's local_stencil (infrastructure/source/operator/
 * r_solver_operator_mod.F90:17-23, operator_real32 of operator_mod.t90).  Either precision may
 * be uploaded to an operator of either storage precision (converted on the device); a download
 * from an FP64 operator rounds to nearest. */
int lfric_op_upload_r32(lfric_op *op, const float *local_stencil);
int lfric_op_download_r32(const lfric_op *op, float *local_stencil);
/* Synthetic SPD element matrices A_e = B_e B_e^T + shift I, B_e ~ N(0,1)/8 (scaled: column j
 * times 10^u_j, u_j ~ U(-2,2)); defined by the counter-based generator (shift 0.1 for the
 * benchmark W0 operators, 1.0 for the W2 mass matrix C C^T + I). */
int lfric_op_generate_spd(lfric_op *op, unsigned long long seed, int scaled, double shift);
/* General element entries scale*U(lo,hi) (optionally times a per-column sign table). */
int lfric_op_generate_uniform(lfric_op *op, unsigned long long seed, unsigned long long stream,
                              double lo, double hi, const double *col_sign);
/* Autotuner of the apply kernels' block order: the 8 XCDs of the GPU (each with its own L2)
 * take workgroups in turn; a window K gives each XCD K consecutive blocks of every 8K.  Times
 * the apply of `form` (0 colour-ordered scatter, 1 gather) for K in {0 (dispatch order), 1,
 * 2, ..., 256} and keeps the fastest in the operator (*best_window, *best_ms per apply).
 * No reference counterpart. */
int lfric_op_tune(lfric_op *op, int form, int *best_window, double *best_ms);
/* Measurement helper (no reference counterpart): `reps` gather-form applies y = A x timed with
 * HIP events on the library stream; phase_ms[0], phase_ms[1] = the average milliseconds of the
 * two kernels of the symmetric-packed apply (per-cell outputs, colour-ordered gather), or of the
 * one-kernel apply and 0. */
int lfric_op_time_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y, int reps, double *phase_ms);
int lfric_op_set_window(lfric_op *op, int window);
/* Add a row-gathered copy of a square lowest-order W0 operator: the rows that feed each
 * output dof stored contiguously, so the gather apply streams the matrix in order.  Costs a
 * second copy of the operator's memory; the gather apply uses it from then on (same sums in
 * the same order).  Writing the operator (upload, generate, convert, invert, assemble)
 * rebuilds the copy. */
int lfric_op_prepare_rows(lfric_op *op);
/* Add a symmetric-packed copy of a square W0 operator of order 0 or 1 whose every element
 * matrix is symmetric bit for bit (refused otherwise): the upper triangle of each element (36
 * of 64 entries at order 0, 378 of 729 at order 1).  Order 0, levels_per_thread 0: one triangle
 * per (cell, level) unit, the previous layer's top rows passed between neighbouring threads;
 * S > 0: segments of S levels per thread, each repeating the previous segment's last triangle.
 * Order 1: one triangle per (cell, layer) unit (levels_per_thread must be 0), the layer below's
 * top rows passed between neighbouring threads, a unit's 18 outputs kept for the second phase.  From then on the gather apply (lfric_lma_apply_gather,
 * the solvers' apply, lfric_lma_apply_mrhs with LFRIC_MRHS_GATHER or AUTO, the batch) reads it
 * in two phases (per-unit row sums, then a colour-ordered gather) whose result equals the
 * colour-ordered scatter (lfric_lma_apply, LFRIC_MRHS_COLOUR) bit for bit.  Writing the
 * operator rebuilds the copy, or drops it if the new entries are not symmetric.  The two
 * phases pass the per-unit outputs through one buffer per operator (for several right-hand
 * sides grown on first use): applies of one operator that overlap on two streams race.  No
 * reference counterpart (the reference stores full element matrices). */
int lfric_op_prepare_symmetric(lfric_op *op, int levels_per_thread);
/* Round an operator into another precision (e.g. the r_solver FP32 copy of an FP64 master). */
int lfric_op_convert(const lfric_op *src, lfric_op *dst);
/* invert_local_operator_code (components/science/source/kernel/algebra/
 * sci_invert_local_operator_kernel_mod.F90:51-77, LU of infrastructure/source/utilities/
 * matrix_invert_mod.F90:62-216): Ainv = per-cell inverse of every element matrix of A, an
 * operator from a space to itself; Ainv on the same spaces and precision, not A itself.  The
 * factorization does not pivot, as the reference's does not. */
int lfric_op_invert_local(const lfric_op *A, lfric_op *Ainv);
/* compute_mass_matrix_w_scalar_code (components/science/source/kernel/fem/
 * sci_compute_mass_matrix_kernel_w_scalar_mod.F90:290-384) for a square lowest-order W0 FP64
 * operator on the device: every cell's mass matrix from the Cartesian coordinate fields chi1..3
 * (FP64 W0 fields on the operator's space, coord_system_xyz, trilinear coordinates), with
 * nqp (1..4) Gauss points per direction (3 integrates a trilinear cell exactly). */
int lfric_op_assemble_mass_w0(lfric_op *op, const lfric_vec *chi1, const lfric_vec *chi2, const lfric_vec *chi3,
                              int nqp);
/* The lowest-order mass matrix of the operator's space, FP64, on the device, from Cartesian
 * coordinate fields chi1..3 (FP64 fields on one lowest-order W0 space of the same mesh and
 * colouring, coord_system_xyz), nqp (1..4) Gauss points per direction:
 *   W0 (8 dofs)  as lfric_op_assemble_mass_w0 (chi on a space with the operator's dofmap);
 *   W2 (6 dofs)  compute_mass_matrix_w2_code (components/science/source/kernel/fem/
 *                sci_compute_mass_matrix_kernel_w2_mod.F90:83-164), Piola-mapped face basis;
 *   W3 (1 dof)   compute_mass_matrix_w3_code (sci_compute_mass_matrix_kernel_w3_mod.F90:270-356),
 *                rehabilitate selecting dj (nonzero, the reference's default) or 1/dj;
 *   Wtheta (2)   compute_mass_matrix_w_scalar_code with the Wtheta basis (bottom 1 - z, top z).
 * The space is recognised from its dofmap's column roles; anything else is refused. */
int lfric_op_assemble_mass(lfric_op *op, const lfric_vec *chi1, const lfric_vec *chi2, const lfric_vec *chi3, int nqp,
                           int rehabilitate);
/* compute_div_operator_code (components/science/source/kernel/fem/
 * sci_compute_div_operator_kernel_mod.F90:120-150): the weak divergence of an FP64 operator
 * from a lowest-order W2 space to a W3 space of the same mesh, from the same coordinate fields;
 * rehabilitate as above. */
int lfric_op_assemble_div(lfric_op *op, const lfric_vec *chi1, const lfric_vec *chi2, const lfric_vec *chi3, int nqp,
                          int rehabilitate);
void lfric_op_destroy(lfric_op *op);

/* invoke(setval_c(y,0), matrix_vector_kernel_type(y, x, A)): y = A x over the coloured
 * mesh with a colour-ordered, atomic-free scatter.  Replaces the whole invoke of
 * components/science/source/kernel/algebra/matrix_vector_kernel_mod.F90:68-140 as called by
 * components/science/source/algorithm/solver/sci_mass_matrix_operator_alg_mod.x90:98-100. */
int lfric_lma_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y);
/* Same product in gather form: one launch; every output dof is written once by the thread of
 * its column, which adds the contributions of the column's cells in colour order, top-partner
 * row of layer k-1 before bottom row of layer k (the order of the reference's coloured loop). */
int lfric_lma_apply_gather(const lfric_op *op, const lfric_vec *x, lfric_vec *y);
/* Several right-hand sides at once: x and y hold nrhs (1, 2, 4 or 8) values per dof, RHS
 * fastest (entry (dof, r) at dof*nrhs + r).  Each element matrix is read once for all
 * right-hand sides; column r of y equals the one-RHS apply of column r bit for bit for the
 * same variant.  No reference counterpart (the reference applies one field at a time). */
#define LFRIC_MRHS_AUTO 0   /* the library's choice (by measured speed: colour from ~200k cells, else gather) */
#define LFRIC_MRHS_GATHER 1 /* one launch, gather form */
#define LFRIC_MRHS_COLOUR 2 /* colour-ordered scatter, one launch per colour */
/* colour-ordered scatter on the matrix cores (v_mfma_f32_4x4x1f32), square lowest-order W0
 * FP32 operators only; bit-identical to LFRIC_MRHS_COLOUR */
#define LFRIC_MRHS_MFMA 3
int lfric_lma_apply_mrhs(const lfric_op *op, const lfric_vec *x, lfric_vec *y, int variant);
/* transpose_matrix_vector_kernel (components/science/source/kernel/algebra/
 * transpose_matrix_vector_kernel_mod.F90): y = A^T x, A maps `from` -> `to`, y on `from`. */
/* lfric_lma_apply on the caller's stream (a hipStream_t; NULL: the library's), returning
 * without waiting: the invoke of setval_c + matrix_vector_kernel_type
 * (sci_mass_matrix_operator_alg_mod.x90:98-100) queued behind the caller's earlier work. */
int lfric_lma_apply_stream(const lfric_op *op, const lfric_vec *x, lfric_vec *y, void *stream);
int lfric_lma_apply_transpose(const lfric_op *op, const lfric_vec *x, lfric_vec *y);
/* dg_matrix_vector_kernel (components/science/source/kernel/algebra/
 * dg_matrix_vector_kernel_mod.F90): discontinuous output, no colouring. */
int lfric_dg_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y);
/* dg_inc_matrix_vector_kernel (components/science/source/kernel/algebra/
 * dg_inc_matrix_vector_kernel_mod.F90:69-142): y = y + A x on a discontinuous output space
 * (GH_READWRITE, no colouring; nothing is zeroed). */
int lfric_dg_inc_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y);
/* enforce_bc_kernel (components/science/source/kernel/algebra/sci_enforce_bc_kernel_mod.F90:
 * 63-91) over every cell column: the vector's dofs that its space flags in boundary_dofs are
 * multiplied by 0 (bottom level, top layer); the others keep their values. */
int lfric_enforce_bc(lfric_vec *v);
/* mm_diagonal_kernel (components/science/source/kernel/algebra/sci_mm_diagonal_kernel_mod.F90
 * :68-125): assembled diagonal of a square LMA operator. */
int lfric_mm_diagonal(const lfric_op *op, lfric_vec *diag);

/* ---- preconditioners (abstract_preconditioner_type, components/science/source/solver/
 *      sci_preconditioner_mod.f90:19-38) ------------------------------------------------ */
/* diagonal_preconditioner_type (components/science/source/algorithm/solver/
 * sci_diagonal_preconditioner_alg_mod.x90:71-106): y = x / diag. */
int lfric_prec_create_diagonal(const lfric_vec *diag, lfric_prec **p);
/* null_preconditioner_type (sci_null_preconditioner_alg_mod.x90): y = x. */
int lfric_prec_create_identity(lfric_prec **p);
int lfric_prec_apply(const lfric_prec *p, const lfric_vec *x, lfric_vec *y);
void lfric_prec_destroy(lfric_prec *p);

/* ---- abstract linear operators (abstract_linear_operator_type, components/science/source/
 *      solver/sci_linear_operator_mod.f90:18-36) ------------------------------------------- */
/* y = A x for a square LMA operator (mass_matrix_operator_type,
 * components/science/source/algorithm/solver/sci_mass_matrix_operator_alg_mod.x90:72-119). */
int lfric_linop_create_lma(const lfric_op *A, lfric_linop **L);
/* mass_matrix_operator_type(mass_matrix, apply_bc) (sci_mass_matrix_operator_alg_mod.x90:42-63,
 * apply :72-119): y = A x, then enforce_bc on y when apply_bc is set (the mass-matrix solver
 * sets it for W1/W2/W2v, sci_mass_matrix_solver_alg_mod.x90:186-191). */
int lfric_linop_create_mass_matrix(const lfric_op *A, int apply_bc, lfric_linop **L);
/* Helmholtz-like reduction of a W2/W3 saddle system on W3:
 *   y = M3 x + tau * D (diag(M2)^-1 (D^T x)),
 * D: W2 -> W3 LMA operator (dg_matrix_vector / transpose_matrix_vector kernels),
 * m2_diag: assembled diagonal of the W2 mass matrix (mm_diagonal), M3: W3 -> W3 LMA operator.
 * All three in one precision.  For lowest-order W2/W3 on one mesh the first apply builds
 * copies of D (transpose-gather order) and of D's rows with M3 (compacted per W3 dof), held by
 * L and rebuilt when D's or M3's entries are rewritten; L must not outlive M3, D or m2_diag. */
int lfric_linop_create_helmholtz(const lfric_op *M3, const lfric_op *D, const lfric_vec *m2_diag,
                                 double tau, lfric_linop **L);
int lfric_linop_apply(const lfric_linop *L, const lfric_vec *x, lfric_vec *y);
void lfric_linop_destroy(lfric_linop *L);

/* Vertical-column block-Jacobi preconditioner: the tridiagonal block of every dof column of the
 * assembled operator (couplings between levels m and m+1 of one column), factorised in FP32
 * and stored in `storage` precision (LFRIC_BF16, LFRIC_F16 or LFRIC_R32); applied with FP32
 * arithmetic to FP32 or FP64 vectors.  Column solves of the reference's
 * components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90 form, built from the
 * element matrices instead of a columnwise operator. */
int lfric_prec_create_column_jacobi(const lfric_linop *L, int storage, lfric_prec **p);

/* ---- iterative solvers (abstract_iterative_solver_type and its CG / GMRES extensions,
 *      components/science/source/solver/sci_iterative_solver_mod.f90:38-57) ----------- */
typedef struct lfric_solver_opts {
  double r_tol;              /* relative tolerance */
  double a_tol;              /* absolute tolerance */
  int max_iter;              /* CG iterations / GMRES restart cycles */
  int gcrk;                  /* GMRES Krylov dimension */
  int monitor_convergence;   /* test convergence (reference: only when true) */
  int fail_on_non_converged; /* return LFRIC_NOT_CONVERGED instead of LFRIC_OK */
  int check_every;           /* host polls the device convergence flag every n iterations */
  double inner_r_tol;        /* mixed precision: inner (FP32) relative reduction per cycle */
  int max_outer;             /* mixed precision: maximum refinement cycles */
  double rho_relax;          /* Jacobi relaxation (solver_config jacobi_relaxation) */
  double lmin, lmax;         /* Chebyshev eigenvalue bounds of P^-1 A (fs_eigenvalues_mod) */
} lfric_solver_opts;

typedef struct lfric_solver_result {
  int status;           /* LFRIC_OK / LFRIC_NOT_CONVERGED / LFRIC_NAN / LFRIC_BREAKDOWN */
  int iterations;       /* total (inner) iterations */
  int outer_iterations; /* refinement cycles (mixed precision), GMRES restart cycles */
  double r_nrm_0;       /* initial residual norm */
  double r_nrm;         /* final residual norm the stopping test used */
  double true_rel_res;  /* ||b - A x|| / ||b - A x0|| in FP64 on exit (mixed and plain) */
} lfric_solver_result;

void lfric_solver_opts_default(lfric_solver_opts *o);

/* ||b - A x|| in FP64-accumulated arithmetic of A's precision (verification helper). */
int lfric_residual_norm(const lfric_op *A, const lfric_vec *x, const lfric_vec *b, double *nrm);

/* cg_solve (sci_iterative_solver_mod.f90:495-592), device resident: x and b are FP64 or
 * FP32 vectors of A's precision; scalars in FP64 on the device. */
int lfric_cg_solve(const lfric_op *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                   const lfric_solver_opts *opts, lfric_solver_result *res);
/* lfric_cg_solve without device allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same kernels and result as lfric_cg_solve; its r, z, p vectors and block partials
 * live in a static 64 MiB device array per GPU (3 vectors of undf values of A's precision plus
 * the partials: undf up to ~2.79M in FP64, ~5.59M in FP32), LFRIC_E_UNSUPPORTED beyond that.
 * The space's gather lists must exist already (built by the first lfric_lma_apply_gather,
 * lfric_op_prepare_rows or solve on the operator), else LFRIC_E_STATE.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int lfric_cg_solve_noalloc(const lfric_op *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                           const lfric_solver_opts *opts, lfric_solver_result *res);
/* FP64 iterative refinement around FP32 (r_solver) CG: r = b - A64 x in FP64, inner CG on
 * A32 with FP32 vectors and FP64 dots, x += d in FP64.  No reference counterpart as one
 * call; the inner solve is cg_solve on r_solver_field_vector_type
 * (components/science/source/algorithm/solver/sci_r_solver_field_vector_mod.x90:119-252). */
int lfric_ir_cg_solve(const lfric_op *A64, const lfric_op *A32, const lfric_prec *P32,
                      lfric_vec *x, const lfric_vec *b, const lfric_solver_opts *opts,
                      lfric_solver_result *res);
/* gmres_solve (sci_iterative_solver_mod.f90:830-1003): left-preconditioned restarted
 * GMRES(gcrk), modified Gram-Schmidt Arnoldi with FP64-accumulated inner products, FP64
 * Hessenberg / Givens least-squares on the device, true-residual test after each cycle.
 * max_iter counts restart cycles, as in the reference. */
int lfric_gmres_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                      const lfric_solver_opts *opts, lfric_solver_result *res);
/* FP64 iterative refinement around FP32 GMRES: r = b - A64 x in FP64, inner GMRES on A32 with
 * FP32 vectors until the residual falls by inner_r_tol, x += d in FP64.  The inner solve reads
 * the Givens least-squares residual estimate after every Arnoldi column (one column behind the
 * device) and ends its cycle once the estimate is below the inner tolerance; res->iterations
 * counts the Arnoldi columns made. */
int lfric_ir_gmres_solve(const lfric_linop *A64, const lfric_linop *A32, const lfric_prec *P32,
                         lfric_vec *x, const lfric_vec *b, const lfric_solver_opts *opts,
                         lfric_solver_result *res);

/* fgmres_solve (sci_iterative_solver_mod.f90:1055-1237): flexible restarted GMRES with a
 * right preconditioner; as in the reference the residual starts from b (zero initial guess
 * assumed) and x += dx at the end.  max_iter counts restart cycles. */
int lfric_fgmres_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                       const lfric_solver_opts *opts, lfric_solver_result *res);
/* bicgstab_solve (sci_iterative_solver_mod.f90:645-779), device resident; the preconditioner
 * acts on the search directions. */
int lfric_bicgstab_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                         const lfric_solver_opts *opts, lfric_solver_result *res);
/* gcr_solve (sci_iterative_solver_mod.f90:1288-1408): restarted GCR(gcrk), right
 * preconditioner, dx from P^-1 b, relative to ||b||; max_iter counts restart cycles. */
int lfric_gcr_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                    const lfric_solver_opts *opts, lfric_solver_result *res);
/* jacobi_solve (sci_iterative_solver_mod.f90, jacobi_smod): x -= rho_relax P^-1 (A x - b). */
int lfric_jacobi_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                       const lfric_solver_opts *opts, lfric_solver_result *res);
/* chebyshev_solve (sci_iterative_solver_mod.f90, chebyshev_smod): max_iter steps from x = 0
 * for eigenvalue bounds [lmin, lmax]; r_nrm = ||A x - b||, r_nrm_0 = max(1, ||b||). */
int lfric_chebyshev_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                          const lfric_solver_opts *opts, lfric_solver_result *res);
/* precondition_only_solve (sci_iterative_solver_mod.f90, precondition_only_smod): x = P^-1 b. */
int lfric_prec_only_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                          const lfric_solver_opts *opts, lfric_solver_result *res);

/* Solver methods of the mass-matrix solve factory (solver_config method_*). */
#define LFRIC_METHOD_CG 1
#define LFRIC_METHOD_BICGSTAB 2
#define LFRIC_METHOD_GMRES 3
#define LFRIC_METHOD_FGMRES 4
#define LFRIC_METHOD_GCR 5
#define LFRIC_METHOD_PREC_ONLY 6
#define LFRIC_METHOD_JACOBI 7
#define LFRIC_METHOD_CHEBYSHEV 8
/* select case(method) of mass_matrix_solver_alg
 * (components/science/source/algorithm/solver/sci_mass_matrix_solver_alg_mod.x90:164-181):
 * one entry point for every method, on an operator and a preconditioner (identity or
 * diagonal in the reference's factory). */
int lfric_solve(int method, const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                const lfric_solver_opts *opts, lfric_solver_result *res);

/* ---- columnwise (CMA) operators -----------------------------------------------------------
 * columnwise_operator_type (infrastructure/source/operator/columnwise_operator_mod.f90) between
 * horizontally discontinuous lowest-order spaces (W3, Wtheta-like: bottom faces, interior, top
 * faces; every column owned by one cell): one banded matrix per cell column with the
 * reference's alpha, beta, gamma_m, gamma_p, bandwidth, nrow, ncol (divided by gcd(alpha,
 * beta)).  Kernels run one thread per column and equal the reference's CMA kernels bit for bit.
 * params[7] = {alpha, beta, gamma_m, gamma_p, bandwidth, nrow, ncol}. */
int lfric_cma_create(const lfric_fs *to, const lfric_fs *from, int precision, lfric_cma **A);
/* columnwise_operator_initialiser_custom: the caller's alpha, beta, gamma_m, gamma_p. */
int lfric_cma_create_custom(const lfric_fs *to, const lfric_fs *from, int precision, int alpha, int beta, int gamma_m,
                            int gamma_p, lfric_cma **A);
int lfric_cma_info(const lfric_cma *A, int *params);
/* columnwise_matrix(bandwidth, nrow, ncell_2d) in the reference layout (Fortran order, FP64). */
int lfric_cma_upload(lfric_cma *A, const double *ref);
int lfric_cma_download(const lfric_cma *A, double *ref);
/* columnwise_op_asm_kernel_code (components/science/source/kernel/columnwise_operators/
 * sci_columnwise_op_asm_kernel_mod.F90:96-140): A from an LMA operator on the same spaces. */
int lfric_cma_assemble(lfric_cma *A, const lfric_op *lma);
/* columnwise_op_asm_diag_hmht_kernel_code (sci_columnwise_op_asm_diag_hmht_kernel_mod.F90:95-135):
 * the column band of Dh diag(M2h)^-1 Dh^T (Dh: W2h -> the operator's space, M2h on W2h). */
int lfric_cma_assemble_diag_hmht(lfric_cma *A, const lfric_op *Dh, const lfric_op *M2h);
/* columnwise_op_asm_m2v_lumped_inv_krnl_code (sci_columnwise_op_asm_m2v_lumped_inv_krnl_mod.F90:
 * 85-113): band 1 = reciprocal of the assembled diagonal of a square LMA operator. */
int lfric_cma_assemble_m2v_lumped_inv(lfric_cma *A, const lfric_op *M);
/* columnwise_op_app_kernel_code (sci_columnwise_op_app_kernel_mod.F90:107-131): y = A x, or
 * y += A x with accumulate != 0 (the kernel alone; the reference's invoke zeroes y first). */
int lfric_cma_apply(const lfric_cma *A, const lfric_vec *x, lfric_vec *y, int accumulate);
/* columnwise_op_appinv_kernel_code (sci_columnwise_op_appinv_kernel_mod.F90:107-146): y = A^-1 x
 * for a square tridiagonal (bandwidth 3) A, Thomas algorithm per column. */
int lfric_cma_apply_inverse(const lfric_cma *A, const lfric_vec *x, lfric_vec *y);
/* columnwise_op_mul_kernel_code (sci_columnwise_op_mul_kernel_mod.F90:128-160): C += A B. */
int lfric_cma_multiply(const lfric_cma *A, const lfric_cma *B, lfric_cma *C);
/* columnwise_op_scaledadd_kernel_code (sci_columnwise_op_scaledadd_kernel_mod.F90:132-166):
 * C = alpha A + beta B (alpha, beta rounded to the operators' precision). */
int lfric_cma_scaled_add(const lfric_cma *A, const lfric_cma *B, lfric_cma *C, double alpha, double beta);
void lfric_cma_destroy(lfric_cma *A);
/* A preconditioner whose apply is lfric_cma_apply_inverse (the vertical line solve of a
 * tridiagonal columnwise operator); A must outlive it. */
int lfric_prec_create_cma_inverse(const lfric_cma *A, lfric_prec **P);
/* tri_solve_code (components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90:57-130):
 * y = T^-1 x on every W3 column, T with diagonal t0, super-diagonal tp and sub-diagonal tm (W3
 * fields); work is a W3 field of the same precision used as scratch. */
int lfric_tri_solve(const lfric_vec *x, const lfric_vec *t0, const lfric_vec *tp, const lfric_vec *tm, lfric_vec *y,
                    lfric_vec *work);

/* ---- batched solves (no reference counterpart: many independent mass-matrix solves at
 *      once) ------------------------------------------------------------------------------
 * Each problem is A x = b with FP64 and FP32 copies of one square LMA operator (ndf 8 or 27),
 * a Jacobi diagonal (FP32) and x, b with 1, 2, 4 or 8 right-hand sides.  Every (problem,
 * RHS) runs FP32 CG inside FP64 iterative refinement (as lfric_ir_cg_solve) with the
 * options given at creation; all problems advance together, one persistent apply kernel and
 * a handful of vector kernels per inner iteration, replayed from a HIP graph.  Problems may
 * be added between runs (continuous batching) and removed once finished.  x is read as the
 * initial guess and holds the solution; b, the operators and the diagonal must stay alive
 * until the problem is removed.  A batch uses the library's stream; calls on one batch must
 * not overlap. */
typedef struct lfric_batch lfric_batch;
int lfric_batch_create(const lfric_solver_opts *opts, lfric_batch **b);
int lfric_batch_add(lfric_batch *b, const lfric_op *A64, const lfric_op *A32, const lfric_vec *diag32,
                    lfric_vec *x, const lfric_vec *rhs, int *id);
/* Advance by up to `rounds` rounds (check_every inner iterations each, then the refinement
 * steps due); *nactive = problems not yet finished. */
int lfric_batch_run(lfric_batch *b, int rounds, int *nactive);
int lfric_batch_solve(lfric_batch *b); /* run until every problem has finished */
int lfric_batch_finished(const lfric_batch *b, int id, int *finished);
/* One result per RHS (res[nrhs]): status, inner iterations, refinement steps, FP64 norms. */
int lfric_batch_result(const lfric_batch *b, int id, lfric_solver_result *res);
int lfric_batch_remove(lfric_batch *b, int id);
/* Autotuner: time the persistent apply of the active problems for each block and grid size
 * and keep the fastest (*best_ms: its time per apply). */
int lfric_batch_tune(lfric_batch *b, double *best_ms);
/* use_graph 0/1 (-1 keeps), block 64/128/256 (0 keeps), grid_mult > 0 (0 keeps). */
int lfric_batch_config(lfric_batch *b, int use_graph, int block, int grid_mult);
/* Multi-RHS apply of the batch: 2 (default) an FP32 operator's copy where it has one (the
 * gather over the row-gathered copy of lowest-order W0, lfric_op_prepare_rows; the two phases
 * of the symmetric-packed order-1 W0 copy, lfric_op_prepare_symmetric), else as 1; 1 colour-
 * ordered where supported (FP32 W0 problems with at most 8 colours: lowest order on the matrix
 * cores as LFRIC_MRHS_MFMA, order 1 one (cell, level) unit per thread as LFRIC_MRHS_COLOUR),
 * gather form for the rest; 0 gather form for every problem.  The symmetric-packed order-1
 * apply equals the colour-ordered one bit for bit, the row-gathered gather the gather form. */
int lfric_batch_set_apply(lfric_batch *b, int form);
int lfric_batch_stats(const lfric_batch *b, long long *iterations, int *block, int *grid_mult, int *ntiles);
void lfric_batch_destroy(lfric_batch *b);

#ifdef __cplusplus
}
#endif
#endif
