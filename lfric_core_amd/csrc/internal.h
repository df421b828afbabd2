// Internal structures of liblfric_hip (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lfric_hip.h"

namespace lfric {

// ---- errors -----------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define LF_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::lfric::fail(LFRIC_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LF_CHECK(cond, code, msg)                   \
  do {                                              \
    if (!(cond)) return ::lfric::fail((code), (msg)); \
  } while (0)

// ---- device context ---------------------------------------------------------------------
struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // scratch for reductions (partials) and a pinned host mirror for scalars
  double *d_partials = nullptr;
  size_t partials_cap = 0;
  unsigned int *d_counter = nullptr;
  double *h_pinned = nullptr;
  long long bytes = 0;
  int apply_form = 1;  // LMA apply inside the solvers: 0 colour-ordered scatter, 1 gather (+ fused dot)
};
Context &ctx();
// the stream the calls of this host thread launch on: the one lfric_set_stream gave, else the
// library's own
hipStream_t cur_stream();
int require_init();
int dmalloc(void **p, size_t bytes);
void dfree(void *p, size_t bytes);

// ---- mesh topology (host) ---------------------------------------------------------------
// Serial-partition extruded cubed-sphere base-layer connectivity.
struct Topology {
  int edge_cells = 0;
  int ncells = 0;
  int nverts = 0, nedges = 0;             // unique 2D vertices / edges (= side faces)
  std::vector<int> cell_next;             // [cell][4] W,S,E,N (1-based)
  std::vector<int> vert_on_cell;          // [cell][4] SW,SE,NE,NW local vertex ids (1-based)
  std::vector<int> face_on_cell;          // [cell][6] W,S,E,N,B,T
  std::vector<int> edge_on_cell;          // [cell][12]
  std::vector<int> colour_of_cell;        // 1..ncolours
  int ncolours = 0;
  std::vector<double> vert_lon, vert_lat; // by local vertex id, radians (generated cubed spheres)
};
int build_cubedsphere_topology(int edge_cells, Topology &t);
// the topology below a global mesh given as [cell][W,S,E,N] neighbours (0: none) and
// [cell][SW,SE,NE,NW] global vertex ids, all 1-based; npanels as the mesh file states it
int build_topology(int ncells, const std::vector<int> &cell_next, const std::vector<int> &vert_global, int npanels,
                   Topology &t);
// UGRID 2D mesh (netCDF classic file): face-node and face-face connectivity, 1-based
int read_ugrid_mesh(const std::string &path, const std::string &mesh_name, std::vector<int> &face_nodes,
                    std::vector<int> &face_links, int &nfaces, int &npanels);
int build_dofmap(const Topology &t, int nlayers, int space, int order, std::vector<int> &dofmap,
                 int &ndf, int &undf);
int extrusion_eta(int kind, int nlayers, std::vector<double> &eta);
struct FunctionSpace;
int cubedsphere_chi(const FunctionSpace &f, int kind, double domain_top, double scaled_radius, double *chi1,
                    double *chi2, double *chi3);

// ---- function space -------------------------------------------------------------------------
struct FunctionSpace {
  int ncells = 0, nlayers = 0, ndf = 0, undf = 0, ncolours = 0, maxcc = 0;
  int space = -1, order = -1;
  std::shared_ptr<Topology> topo;          // set for generated cubed-sphere spaces
  std::vector<int> dofmap;                 // [cell][ndf] 1-based (reference layout)
  std::vector<int> npc;                    // cells per colour
  std::vector<int> cic;                    // [colour][maxcc] 1-based cells
  std::vector<int> colour_start;           // prefix sums of npc, size ncolours+1
  std::vector<int> cell_of_pos;            // colour-major position -> cell (0-based)
  // column roles: df is "top" when map(df) = map(bottom)+1 in every cell
  uint32_t top_mask = 0;
  std::vector<int> partner;                // bottom df -> top df or -1
  // dof_on_vert_boundary (function_space_mod.F90:791-801): bit df set when boundary_value(df,1)
  // (bottom level) or boundary_value(df,2) (top layer) is 0
  uint32_t bc_bottom = 0, bc_top = 0;
  // operator element slots: column pos owns ebase[pos] + 0..nlayers; colours start 64-aligned
  std::vector<int> ebase;
  long long nelem_padded = 0;
  // device
  int *d_ebase = nullptr;
  int *d_map = nullptr;                    // [pos][ndf] 0-based column-base dofs
  uint32_t *d_first = nullptr;             // [pos] bit df: first touch of df's column
  uint32_t *d_last = nullptr;              // [pos] bit df: last touch of df's column
  size_t map_bytes = 0, flag_bytes = 0, ebase_bytes = 0;
  // gather form of an apply into this space (built on first use): one list per dof column
  // (sorted by base dof) of the (cell position, bottom-role df) pairs that write it, in colour
  // order
  int gather_ncols = 0;
  int2 *d_gcol = nullptr;                  // [col] base dof (0-based), 1 when the column is nl+1 high
  int *d_goff = nullptr;                   // [col+1] offsets into d_gent
  int *d_gent = nullptr;                   // pos * 32 + df
  size_t gcol_bytes = 0, goff_bytes = 0, gent_bytes = 0;
  ~FunctionSpace();
};

}  // namespace lfric

struct lfric_fs {
  std::shared_ptr<lfric::FunctionSpace> fs;
};

struct lfric_vec {
  std::shared_ptr<lfric::FunctionSpace> fs;
  int precision = LFRIC_R64;
  int nrhs = 1;
  long long n = 0;  // undf * nrhs
  void *d = nullptr;
  size_t bytes = 0;
};

struct lfric_op {
  std::shared_ptr<lfric::FunctionSpace> to, from;
  int precision = LFRIC_R64;
  size_t gs = 0;               // values per 64-element group: 64 * (ndf_to*ndf_from padded to 16-byte chunks)
  long long ngroups = 0;       // 64-element groups
  uint32_t row_top = 0;        // rows stored one level up (top dofs of the to space)
  int xwin = 0;                // XCD block-order window of the apply kernels (lfric_op_tune)
  // optional row-gathered copy for the gather apply (lfric_op_prepare_rows): the matrix rows
  // that feed each output dof, stored contiguously per thread of the gather kernel
  void *rows = nullptr;
  size_t rows_bytes = 0, gsr = 0;
  // symmetric-packed copy (lfric_op_prepare_symmetric): upper triangles per (cell, level
  // segment); per-unit outputs in symz between the two phases of its apply
  void *sym = nullptr, *symz = nullptr;
  size_t sym_bytes = 0, symz_bytes = 0, gss = 0;
  int sym_seg = 0;
  int4 *d_symzu = nullptr;     // per gather column: its cells' unit ids (cell*4 + i), -1 beyond them
  size_t symzu_bytes = 0;
  int4 *d_symcol = nullptr;    // one-unit-per-thread form: the k-th cell's 4 bottom vertex column bases
  size_t symcol_bytes = 0;
  // order-1 W0 copy: symz holds one right-hand side's per-unit row sums, d_symzu the gather
  // columns' cells as pos*18 + (slot of dof i), d_sympos the k-th cell's colour-major position; symz_mr the
  // row sums of several right-hand sides, grown on first use
  int *d_sympos = nullptr;
  size_t sympos_bytes = 0;
  mutable void *symz_mr = nullptr;
  mutable size_t symz_mr_bytes = 0;
  void *d = nullptr;           // level-aligned AoSoA layout (see kernels.hip)
  size_t bytes = 0;
  unsigned long long version = 0;  // counts the rewrites of the entries (copies held elsewhere follow it)
};

// Columnwise (CMA) operator between horizontally discontinuous spaces: one banded matrix per
// cell column, entry (row i, band b) of column position pos at d[(i*bw + b)*ncols + pos] so that
// the threads of a wavefront (one column each) read consecutive addresses.
struct lfric_cma {
  std::shared_ptr<lfric::FunctionSpace> to, from;
  int precision = LFRIC_R32;
  int alpha = 0, beta = 0, gamma_m = 0, gamma_p = 0, bw = 0, nrow = 0, ncol = 0;
  int ni_to = 0, nf_to = 0, ni_from = 0, nf_from = 0;
  int ncols = 0;                       // cell columns (positions of the to space)
  void *d = nullptr;
  size_t bytes = 0;
  int *d_ind_to = nullptr, *d_ind_from = nullptr;        // row/column -> offset from the column base
  int *d_band_to = nullptr, *d_band_from = nullptr;      // [k][df] banded row/column (1-based)
  size_t ind_to_bytes = 0, ind_from_bytes = 0, band_to_bytes = 0, band_from_bytes = 0;
  // [nrow][ncols] forward-sweep factors of the tridiagonal solve: scratch grown on first use,
  // not part of the operator's value
  mutable void *work = nullptr;
  mutable size_t work_bytes = 0;
  ~lfric_cma();
};

struct lfric_prec {
  int kind = 0;  // 0 identity, 1 diagonal, 2 vertical-column block-Jacobi, 3 columnwise tridiagonal inverse
  const lfric_vec *diag = nullptr;
  const lfric_cma *cma = nullptr;
  // column block-Jacobi: columns (base dof, height) and per-dof Thomas factors
  int storage = LFRIC_R32;
  long long undf = 0;
  int ncols = 0;
  int *d_cols = nullptr;   // [ncols][2] base (0-based), height; sorted by base
  int tile = 0, cap = 0;   // columns per block of the LDS-tiled apply (0: not used), dofs per tile at most
  void *d_fac = nullptr;   // [3][undf]: lower multiplier, 1/pivot, super-diagonal
  size_t cols_bytes = 0, fac_bytes = 0;
  ~lfric_prec();
};

struct lfric_linop {
  int kind = 0;  // 0 LMA, 1 Helmholtz composite
  int apply_bc = 0;  // kind 0: enforce the space's boundary conditions after each apply
  int precision = LFRIC_R64;
  const lfric_op *A = nullptr;
  const lfric_op *M3 = nullptr, *D = nullptr;
  const lfric_vec *m2 = nullptr;
  double tau = 0.0;
  std::shared_ptr<lfric::FunctionSpace> space;  // domain and range
  void *t_w2 = nullptr, *t_w3 = nullptr;      // scratch of the composite apply
  size_t t_w2_bytes = 0, t_w3_bytes = 0;
  mutable int fused = -1;  // kind 1: the two-kernel composite applies (set on first use)
  mutable int hcol_state = 0;  // kind 1: 1 the W2 column table is built, -1 the mesh does not allow one
  mutable void *d_hcol = nullptr;
  mutable size_t hcol_bytes = 0;
  // kind 1: D's entries in transpose-gather order (two per W2 dof) and the D version they hold
  mutable void *d_dtc = nullptr;
  mutable size_t dtc_bytes = 0;
  mutable unsigned long long dtc_version = ~0ull;
  // kind 1: D's rows and M3 compacted in the divergence kernel's thread order, and the D and
  // M3 versions they hold
  mutable void *d_drc = nullptr;
  mutable size_t drc_bytes = 0;
  mutable unsigned long long drc_dver = ~0ull, drc_mver = ~0ull;
  ~lfric_linop();
};
