// C ABI of liblfric_hip: handles, set-up and the device-resident solver drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "internal.h"
#include "kernels.h"

namespace lfric {

// ---- errors and context ---------------------------------------------------------------------
static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

static Context g_ctx;
Context &ctx() { return g_ctx; }
static thread_local hipStream_t t_stream = nullptr;
hipStream_t cur_stream() { return t_stream ? t_stream : g_ctx.stream; }

int require_init() {
  if (g_ctx.device < 0) return fail(LFRIC_E_STATE, "lfric_init has not been called");
  return LFRIC_OK;
}

int dmalloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 64;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(LFRIC_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  g_ctx.bytes += static_cast<long long>(bytes);
  return LFRIC_OK;
}

void dfree(void *p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  g_ctx.bytes -= static_cast<long long>(bytes == 0 ? 64 : bytes);
}

FunctionSpace::~FunctionSpace() {
  dfree(d_map, map_bytes);
  dfree(d_first, flag_bytes);
  dfree(d_last, flag_bytes);
  dfree(d_ebase, ebase_bytes);
  dfree(d_gcol, gcol_bytes);
  dfree(d_goff, goff_bytes);
  dfree(d_gent, gent_bytes);
}

// Scratch vectors of the solvers, kept between solves (grown on demand).  Slots: 0-2 CG's r, z,
// p; 3 the true-residual vector; 4 CG's p.Ap partials; 5, 6 and 8 the refinement driver's FP32
// right-hand side, FP32 correction and FP64 residual (live across its inner CG solves); 7 host
// staging.
struct Workspace {
  void *p = nullptr;
  size_t bytes = 0;
};
static Workspace g_ws[9];
static SolverState *g_state = nullptr;
static SolverState *g_state_host = nullptr;

// Device memory of the allocation-free CG (lfric_cg_solve_noalloc): its r, z and p vectors and
// the fused apply's p.Ap block partials, in place of workspace slots 0-2 and 4.  Every GPU the
// library runs on holds its own copy until the process ends.  The array's name is a host-side
// stand-in: host code uses only the address hipGetSymbolAddress returned for the device
// (g_scratch_ptr), recorded by lfric_init with that device current.
constexpr size_t kScratchBytes = size_t{64} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices] = {};

static int ws_get(int slot, size_t bytes, void **out) {
  Workspace &w = g_ws[slot];
  if (w.bytes < bytes) {
    dfree(w.p, w.bytes);
    w.p = nullptr;
    w.bytes = 0;
    if (int rc = dmalloc(&w.p, bytes)) return rc;
    w.bytes = bytes;
  }
  *out = w.p;
  return LFRIC_OK;
}

static size_t esize(int prec) {
  switch (prec) {
    case LFRIC_R64: return 8;
    case LFRIC_R32: return 4;
    case LFRIC_BF16:
    case LFRIC_F16: return 2;
  }
  return 0;
}

// ---- function spaces -----------------------------------------------------------------------
static int finish_space(std::shared_ptr<FunctionSpace> &fs) {
  FunctionSpace &f = *fs;
  const int nc = f.ncells, ndf = f.ndf;
  LF_CHECK(ndf >= 1 && ndf <= 32, LFRIC_E_UNSUPPORTED, "ndf must be in [1, 32]");
  // colour-major positions
  f.colour_start.assign(f.ncolours + 1, 0);
  for (int c = 0; c < f.ncolours; ++c) f.colour_start[c + 1] = f.colour_start[c] + f.npc[c];
  LF_CHECK(f.colour_start[f.ncolours] == nc, LFRIC_E_ARG, "colouring does not cover every cell once");
  f.cell_of_pos.resize(nc);
  std::vector<char> seen_cell(nc, 0);
  for (int c = 0; c < f.ncolours; ++c)
    for (int i = 0; i < f.npc[c]; ++i) {
      int cell = f.cic[static_cast<size_t>(c) * f.maxcc + i];
      LF_CHECK(cell >= 1 && cell <= nc && !seen_cell[cell - 1], LFRIC_E_ARG, "invalid or repeated cell in colouring");
      seen_cell[cell - 1] = 1;
      f.cell_of_pos[f.colour_start[c] + i] = cell - 1;
    }
  for (int i = 0; i < nc * ndf; ++i)
    LF_CHECK(f.dofmap[i] >= 1 && f.dofmap[i] + f.nlayers - 1 <= f.undf, LFRIC_E_ARG, "dofmap entry out of range");
  // column roles.  Two dofs of one cell whose columns overlap must be a bottom/top pair
  // (map(top) = map(bottom) + 1, the dofmap convention of dofmap_setup); the pair set must be
  // the same in every cell.  Non-overlapping columns need no role.
  f.top_mask = 0;
  f.partner.assign(ndf, -1);
  for (int c = 0; c < nc; ++c) {
    const int *row = &f.dofmap[static_cast<size_t>(c) * ndf];
    for (int a = 0; a < ndf; ++a)
      for (int b = 0; b < ndf; ++b) {
        const int d = row[b] - row[a];
        if (a == b || d <= 0 || d >= f.nlayers) continue;
        LF_CHECK(d == 1, LFRIC_E_UNSUPPORTED, "dofmap columns of one cell overlap by more than one level");
        if (c == 0) {
          LF_CHECK(f.partner[a] < 0 && !((f.top_mask >> b) & 1u), LFRIC_E_UNSUPPORTED,
                   "dofmap has a dof with two top partners");
          f.partner[a] = b;
          f.top_mask |= 1u << b;
        } else {
          LF_CHECK(f.partner[a] == b, LFRIC_E_UNSUPPORTED, "bottom/top dof pairs differ between cells");
        }
      }
  }
  for (int a = 0; a < ndf; ++a)
    if (f.partner[a] >= 0 && ((f.top_mask >> a) & 1u))
      return fail(LFRIC_E_UNSUPPORTED, "dofmap has chained top/bottom dofs");
  for (int a = 0; a < ndf; ++a)  // every pair of cell 0 must exist in all cells
    if (f.partner[a] >= 0)
      for (int c = 1; c < nc; ++c)
        LF_CHECK(f.dofmap[static_cast<size_t>(c) * ndf + f.partner[a]] == f.dofmap[static_cast<size_t>(c) * ndf + a] + 1,
                 LFRIC_E_UNSUPPORTED, "bottom/top dof pairs differ between cells");
  // colouring must keep every column to one cell per colour (race-free scatter)
  {
    std::vector<int> stamp(static_cast<size_t>(f.undf) + 2, -1);
    for (int c = 0; c < f.ncolours; ++c)
      for (int p = f.colour_start[c]; p < f.colour_start[c + 1]; ++p) {
        const int cell = f.cell_of_pos[p];
        for (int d = 0; d < ndf; ++d) {
          int &s = stamp[f.dofmap[cell * ndf + d]];
          if (s == c) return fail(LFRIC_E_ARG, "two cells of one colour share a dof column");
          s = c;
        }
      }
  }
  // first / last touch of each bottom-role column in processing order
  std::vector<uint32_t> first(nc, 0), last(nc, 0);
  {
    std::vector<int> last_pos(static_cast<size_t>(f.undf) + 2, -1);
    for (int p = 0; p < nc; ++p) {
      const int cell = f.cell_of_pos[p];
      for (int d = 0; d < ndf; ++d) {
        if ((f.top_mask >> d) & 1u) continue;
        int &lp = last_pos[f.dofmap[cell * ndf + d]];
        if (lp < 0) first[p] |= 1u << d;
        lp = p;
      }
    }
    for (int p = 0; p < nc; ++p) {
      const int cell = f.cell_of_pos[p];
      for (int d = 0; d < ndf; ++d)
        if (!((f.top_mask >> d) & 1u) && last_pos[f.dofmap[cell * ndf + d]] == p) last[p] |= 1u << d;
    }
  }
  std::vector<int> mpos(static_cast<size_t>(nc) * ndf);
  for (int p = 0; p < nc; ++p)
    for (int d = 0; d < ndf; ++d) mpos[static_cast<size_t>(p) * ndf + d] = f.dofmap[f.cell_of_pos[p] * ndf + d] - 1;
  // element slots of the operator layout
  f.ebase.resize(nc);
  {
    long long e0 = 0;
    for (int c = 0; c < f.ncolours; ++c) {
      for (int p = f.colour_start[c]; p < f.colour_start[c + 1]; ++p)
        f.ebase[p] = static_cast<int>(e0 + static_cast<long long>(p - f.colour_start[c]) * (f.nlayers + 1));
      e0 += static_cast<long long>(f.colour_start[c + 1] - f.colour_start[c]) * (f.nlayers + 1);
      e0 = (e0 + 63) / 64 * 64;
    }
    LF_CHECK(e0 < (1LL << 31), LFRIC_E_UNSUPPORTED, "mesh too large for 32-bit element slots");
    f.nelem_padded = std::max<long long>(e0, 64);
  }
  f.ebase_bytes = sizeof(int) * nc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&f.d_ebase), f.ebase_bytes)) return rc;
  LF_HIP(hipMemcpy(f.d_ebase, f.ebase.data(), f.ebase_bytes, hipMemcpyHostToDevice));
  f.map_bytes = sizeof(int) * mpos.size();
  f.flag_bytes = sizeof(uint32_t) * nc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&f.d_map), f.map_bytes)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&f.d_first), f.flag_bytes)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&f.d_last), f.flag_bytes)) return rc;
  LF_HIP(hipMemcpy(f.d_map, mpos.data(), f.map_bytes, hipMemcpyHostToDevice));
  LF_HIP(hipMemcpy(f.d_first, first.data(), f.flag_bytes, hipMemcpyHostToDevice));
  LF_HIP(hipMemcpy(f.d_last, last.data(), f.flag_bytes, hipMemcpyHostToDevice));
  return LFRIC_OK;
}

static int space_from_topology(const std::shared_ptr<Topology> &t, int nlayers, int space, int order,
                               lfric_fs **out) {
  auto fs = std::make_shared<FunctionSpace>();
  FunctionSpace &f = *fs;
  if (int rc = build_dofmap(*t, nlayers, space, order, f.dofmap, f.ndf, f.undf)) return rc;
  f.ncells = t->ncells;
  f.nlayers = nlayers;
  f.space = space;
  f.order = order;
  f.topo = t;
  f.ncolours = t->ncolours;
  f.npc.assign(f.ncolours, 0);
  for (int c = 0; c < t->ncells; ++c) f.npc[t->colour_of_cell[c] - 1]++;
  f.maxcc = *std::max_element(f.npc.begin(), f.npc.end());
  f.cic.assign(static_cast<size_t>(f.ncolours) * f.maxcc, 0);
  std::vector<int> fill(f.ncolours, 0);
  for (int c = 1; c <= t->ncells; ++c) {
    const int col = t->colour_of_cell[c - 1] - 1;
    f.cic[static_cast<size_t>(col) * f.maxcc + fill[col]++] = c;
  }
  // dof_on_vert_boundary of the generated lowest-order spaces
  // (function_space_constructor_helper_functions_mod.F90:678-680, 1048-1049): W2 flags its
  // bottom face (dof 5) at the bottom and its top face (dof 6) at the top; W0 and W3 none
  if (space == LFRIC_W2 && order == 0 && f.ndf == 6) {
    f.bc_bottom = 1u << 4;
    f.bc_top = 1u << 5;
  }
  if (int rc = finish_space(fs)) return rc;
  *out = new lfric_fs{fs};
  return LFRIC_OK;
}

// ---- solver drivers ---------------------------------------------------------------------------
struct CgArgs {
  const lfric_op *A;
  const void *diag;  // nullptr: identity
  void *x;
  const void *b;
  long long n;
  int prec;
  bool x_zero;
  double r_tol, a_tol;
  int max_iter, monitor, check_every;
  bool bc = false;  // enforce the space's boundary conditions after every apply (mass-matrix operator)
};

// cg_solve (components/science/source/solver/sci_iterative_solver_mod.f90:495-592) with every
// vector and scalar on the device.  The host enqueues `check_every` iterations at a time and
// then reads the device state once; kernels of iterations past convergence return at once.
// Byte offsets of r, z, p and the p.Ap partials inside the allocation-free CG's scratch array:
// each piece starts on a 256-byte boundary.  Returns the bytes used.
static size_t cg_scratch_layout(size_t vb, int pz_blocks, size_t off[4]) {
  const auto up = [](size_t v) { return (v + 255) / 256 * 256; };
  off[0] = 0;
  off[1] = up(vb);
  off[2] = off[1] + up(vb);
  off[3] = off[2] + up(vb);
  return off[3] + up(sizeof(double) * static_cast<size_t>(std::max(pz_blocks, 1)));
}

// scratch_base: null for the grown workspace slots; else the device address of the
// allocation-free CG's scratch array, already checked to hold cg_scratch_layout's bytes.
static int cg_device(const CgArgs &a, SolverState &out, unsigned char *scratch_base = nullptr) {
  Context &c = ctx();
  const size_t vb = esize(a.prec) * static_cast<size_t>(a.n);
  void *r, *z, *p;
  size_t soff[4] = {0, 0, 0, 0};
  if (scratch_base) {
    cg_scratch_layout(vb, 1, soff);
    r = scratch_base + soff[0];
    z = scratch_base + soff[1];
    p = scratch_base + soff[2];
  } else {
    if (int rc = ws_get(0, vb, &r)) return rc;
    if (int rc = ws_get(1, vb, &z)) return rc;
    if (int rc = ws_get(2, vb, &p)) return rc;
  }
  SolverState s{};
  s.r_tol = a.r_tol;
  s.a_tol = a.a_tol;
  s.max_iter = a.max_iter;
  s.monitor = a.monitor;
  LF_HIP(hipMemcpyAsync(g_state, &s, sizeof(s), hipMemcpyHostToDevice, cur_stream()));
  if (a.x_zero) {
    if (int rc = launch_copy(a.prec, r, a.b, a.n)) return rc;
  } else {
    if (int rc = launch_lma_apply_solver(a.A, a.x, r, nullptr)) return rc;
    if (a.bc)
      if (int rc = launch_enforce_bc(*a.A->to, a.prec, r, nullptr)) return rc;
    if (int rc = launch_cg_residual(a.prec, r, a.b, a.n)) return rc;
  }
  if (int rc = launch_cg_init_norm(a.prec, r, a.n, g_state)) return rc;
  if (int rc = launch_cg_init_prec(a.prec, r, z, p, a.diag, a.n, g_state)) return rc;
  const int every = std::max(1, a.check_every);
  const int max_chunks = a.max_iter / every + 2;
  // gather form: z = A p and the partial sums of p.z in one kernel
  const bool fused = c.apply_form == 1 && a.A->to == a.A->from && !a.bc;
  double *pz_partials = nullptr;
  int pz_blocks = 0;
  if (fused) {
    if (int rc = launch_gather_blocks(a.A, &pz_blocks)) return rc;
    void *w;
    if (scratch_base) {
      w = scratch_base + soff[3];
    } else if (int rc = ws_get(4, sizeof(double) * static_cast<size_t>(pz_blocks), &w)) {
      return rc;
    }
    pz_partials = static_cast<double *>(w);
  }
  // the p.(A p) partials: the fused apply's, or those of a separate dot (first kRedBlocks
  // partials of the context's buffer); the update adds them itself when there are few
  const double *pzp = fused ? pz_partials : c.d_partials;
  const int npz = fused ? pz_blocks : kRedBlocks;
  const bool fold = npz <= kFoldMaxPartials;
  long long it = 0;  // iterations enqueued: their parity picks the rz slot
  for (int ch = 0; ch < max_chunks; ++ch) {
    for (int i = 0; i < every; ++i, ++it) {
      const int par = static_cast<int>(it & 1);
      if (fused) {
        if (int rc = launch_lma_apply_gather_dot(a.A, p, z, pz_partials, &g_state->done)) return rc;
      } else {
        if (int rc = launch_lma_apply_solver(a.A, p, z, &g_state->done)) return rc;
        if (a.bc)
          if (int rc = launch_enforce_bc(*a.A->to, a.prec, z, &g_state->done)) return rc;
        if (int rc = launch_cg_dot_pz(a.prec, p, z, a.n, g_state, c.d_partials)) return rc;
      }
      if (!fold)
        if (int rc = launch_cg_finish_pz(pzp, npz, g_state, par)) return rc;
      if (int rc = launch_cg_update(a.prec, a.x, r, z, p, a.diag, a.n, g_state, fold ? pzp : nullptr, npz, par))
        return rc;
      if (int rc = launch_cg_aypx(a.prec, p, z, a.n, g_state, par)) return rc;
    }
    LF_HIP(hipMemcpyAsync(g_state_host, g_state, sizeof(SolverState), hipMemcpyDeviceToHost, cur_stream()));
    LF_HIP(hipStreamSynchronize(cur_stream()));
    if (g_state_host->done) break;
  }
  out = *g_state_host;
  return LFRIC_OK;
}

static int true_residual(const lfric_op *A, const void *x, const void *b, long long n, int prec, double *nrm) {
  void *r;
  if (int rc = ws_get(3, esize(prec) * static_cast<size_t>(n), &r)) return rc;
  if (int rc = launch_lma_apply_solver(A, x, r, nullptr)) return rc;
  if (int rc = launch_cg_residual(prec, r, b, n)) return rc;
  double *d_out = &g_state->red0;
  if (int rc = launch_dot(prec, r, r, n, d_out)) return rc;
  double h = 0;
  LF_HIP(hipMemcpyAsync(g_state_host, g_state, sizeof(SolverState), hipMemcpyDeviceToHost, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  h = g_state_host->red0;
  *nrm = std::sqrt(h);
  return LFRIC_OK;
}

}  // namespace lfric

using namespace lfric;

// ======================================================================================
extern "C" {

int lfric_init(int device) {
  Context &c = ctx();
  if (c.device >= 0) {
    if (c.device == device) return LFRIC_OK;
    return fail(LFRIC_E_STATE, "library already initialised on another device");
  }
  int n = 0;
  LF_HIP(hipGetDeviceCount(&n));
  LF_CHECK(device >= 0 && device < n, LFRIC_E_ARG, "no such device");
  LF_HIP(hipSetDevice(device));
  LF_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  LF_HIP(hipEventCreate(&c.t0));
  LF_HIP(hipEventCreate(&c.t1));
  c.device = device;
  c.partials_cap = static_cast<size_t>(kRedBlocks) * 4;
  if (int rc = dmalloc(reinterpret_cast<void **>(&c.d_partials), sizeof(double) * c.partials_cap)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&c.d_counter), sizeof(unsigned int) * 16)) return rc;
  LF_HIP(hipMemset(c.d_counter, 0, sizeof(unsigned int) * 16));
  if (int rc = dmalloc(reinterpret_cast<void **>(&g_state), sizeof(SolverState))) return rc;
  LF_HIP(hipMemset(g_state, 0, sizeof(SolverState)));
  LF_HIP(hipHostMalloc(reinterpret_cast<void **>(&g_state_host), sizeof(SolverState), hipHostMallocDefault));
  // this device's copy of the allocation-free CG's scratch array
  LF_CHECK(device < kMaxDevices, LFRIC_E_UNSUPPORTED, "device index beyond the scratch table");
  void *sp = nullptr;
  LF_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  LF_CHECK(sp, LFRIC_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_ptr[device] = sp;
  return LFRIC_OK;
}

const char *lfric_last_error(void) { return g_err.c_str(); }

int lfric_set_stream(void *stream) {
  if (int rc = require_init()) return rc;
  t_stream = static_cast<hipStream_t>(stream);
  return LFRIC_OK;
}

int lfric_synchronize(void) {
  if (int rc = require_init()) return rc;
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}

int lfric_timer_start(void) {
  if (int rc = require_init()) return rc;
  LF_HIP(hipEventRecord(ctx().t0, cur_stream()));
  return LFRIC_OK;
}

int lfric_timer_stop(double *ms) {
  if (int rc = require_init()) return rc;
  LF_HIP(hipEventRecord(ctx().t1, cur_stream()));
  LF_HIP(hipEventSynchronize(ctx().t1));
  float f = 0.f;
  LF_HIP(hipEventElapsedTime(&f, ctx().t0, ctx().t1));
  *ms = f;
  return LFRIC_OK;
}

long long lfric_device_bytes(void) { return ctx().bytes; }

int lfric_set_apply_form(int form) {
  if (int rc = require_init()) return rc;
  LF_CHECK(form == 0 || form == 1, LFRIC_E_ARG, "apply form must be 0 (colour-ordered scatter) or 1 (gather)");
  ctx().apply_form = form;
  return LFRIC_OK;
}

int lfric_device_memory(long long *free_bytes, long long *total_bytes) {
  if (int rc = require_init()) return rc;
  size_t f = 0, t = 0;
  LF_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = static_cast<long long>(f);
  if (total_bytes) *total_bytes = static_cast<long long>(t);
  return LFRIC_OK;
}

// ---- function spaces ------------------------------------------------------------------------
int lfric_fs_create(int ncells, int nlayers, int ndf, int undf, const int *dofmap, int ncolours,
                    const int *ncells_per_colour, const int *cells_in_colour, int maxcc, lfric_fs **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(out && dofmap && ncells_per_colour && cells_in_colour, LFRIC_E_ARG, "null argument");
  LF_CHECK(ncells > 0 && nlayers > 0 && ndf > 0 && undf > 0 && ncolours > 0 && maxcc > 0, LFRIC_E_ARG,
           "sizes must be positive");
  auto fs = std::make_shared<FunctionSpace>();
  FunctionSpace &f = *fs;
  f.ncells = ncells;
  f.nlayers = nlayers;
  f.ndf = ndf;
  f.undf = undf;
  f.ncolours = ncolours;
  f.maxcc = maxcc;
  f.dofmap.assign(dofmap, dofmap + static_cast<size_t>(ncells) * ndf);
  f.npc.assign(ncells_per_colour, ncells_per_colour + ncolours);
  f.cic.assign(cells_in_colour, cells_in_colour + static_cast<size_t>(ncolours) * maxcc);
  for (int c = 0; c < ncolours; ++c)
    LF_CHECK(f.npc[c] >= 0 && f.npc[c] <= maxcc, LFRIC_E_ARG, "ncells_per_colour exceeds maxcc");
  if (int rc = finish_space(fs)) return rc;
  *out = new lfric_fs{fs};
  return LFRIC_OK;
}

int lfric_fs_create_cubedsphere(int edge_cells, int nlayers, int space, int order, lfric_fs **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(out && nlayers > 0, LFRIC_E_ARG, "bad arguments");
  auto t = std::make_shared<Topology>();
  if (int rc = build_cubedsphere_topology(edge_cells, *t)) return rc;
  return space_from_topology(t, nlayers, space, order, out);
}

int lfric_ugrid_read(const char *path, const char *mesh_name, int *nfaces, int *npanels, int *face_nodes,
                     int *face_links) {
  LF_CHECK(path && nfaces, LFRIC_E_ARG, "null argument");
  std::vector<int> fn, fl;
  int nf = 0, np = 0;
  if (int rc = read_ugrid_mesh(path, mesh_name ? mesh_name : "", fn, fl, nf, np)) return rc;
  *nfaces = nf;
  if (npanels) *npanels = np;
  if (face_nodes) std::copy(fn.begin(), fn.end(), face_nodes);
  if (face_links) std::copy(fl.begin(), fl.end(), face_links);
  return LFRIC_OK;
}

int lfric_fs_create_ugrid(const char *path, const char *mesh_name, int nlayers, int space, int order,
                          lfric_fs **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(path && out && nlayers > 0, LFRIC_E_ARG, "bad arguments");
  std::vector<int> fn, fl;
  int nf = 0, np = 0;
  if (int rc = read_ugrid_mesh(path, mesh_name ? mesh_name : "", fn, fl, nf, np)) return rc;
  auto t = std::make_shared<Topology>();
  if (int rc = build_topology(nf, fl, fn, np, *t)) return rc;
  return space_from_topology(t, nlayers, space, order, out);
}

int lfric_fs_create_sibling(const lfric_fs *like, int space, int order, lfric_fs **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(like && out && like->fs->topo, LFRIC_E_ARG, "sibling spaces need a generated mesh");
  return space_from_topology(like->fs->topo, like->fs->nlayers, space, order, out);
}

int lfric_fs_info(const lfric_fs *fs, int *ncells, int *nlayers, int *ndf, int *undf, int *ncolours) {
  LF_CHECK(fs, LFRIC_E_ARG, "null space");
  const FunctionSpace &f = *fs->fs;
  if (ncells) *ncells = f.ncells;
  if (nlayers) *nlayers = f.nlayers;
  if (ndf) *ndf = f.ndf;
  if (undf) *undf = f.undf;
  if (ncolours) *ncolours = f.ncolours;
  return LFRIC_OK;
}

int lfric_fs_get_dofmap(const lfric_fs *fs, int *dofmap) {
  LF_CHECK(fs && dofmap, LFRIC_E_ARG, "null argument");
  std::copy(fs->fs->dofmap.begin(), fs->fs->dofmap.end(), dofmap);
  return LFRIC_OK;
}

int lfric_fs_get_colouring(const lfric_fs *fs, int *npc, int *cic) {
  LF_CHECK(fs, LFRIC_E_ARG, "null space");
  const FunctionSpace &f = *fs->fs;
  if (npc) std::copy(f.npc.begin(), f.npc.end(), npc);
  if (cic) std::copy(f.cic.begin(), f.cic.end(), cic);
  return LFRIC_OK;
}

void lfric_fs_destroy(lfric_fs *fs) { delete fs; }

// ---- vectors ----------------------------------------------------------------------------------
// Buffers of destroyed vectors kept for the next vector of the same size: a Krylov solver's
// copy() releases and recreates a vector every call, and hipFree waits for the whole device.
// Only buffers up to 64 MiB are kept, at most 256 MiB in all; they count as free, not as held
// by live handles.
static std::mutex g_vec_cache_mu;
static std::vector<std::pair<void *, size_t>> g_vec_cache;
static size_t g_vec_cache_bytes = 0;
static constexpr size_t kVecCacheEntryMax = size_t(64) << 20, kVecCacheTotalMax = size_t(256) << 20;

static int vec_buffer_get(void **p, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(g_vec_cache_mu);
    for (size_t i = 0; i < g_vec_cache.size(); ++i)
      if (g_vec_cache[i].second == bytes) {
        *p = g_vec_cache[i].first;
        g_vec_cache.erase(g_vec_cache.begin() + static_cast<long>(i));
        g_vec_cache_bytes -= bytes;
        g_ctx.bytes += static_cast<long long>(bytes);
        return LFRIC_OK;
      }
  }
  return dmalloc(p, bytes);
}

static void vec_buffer_put(void *p, size_t bytes) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_vec_cache_mu);
    if (bytes > 0 && bytes <= kVecCacheEntryMax && g_vec_cache_bytes + bytes <= kVecCacheTotalMax) {
      g_vec_cache.emplace_back(p, bytes);
      g_vec_cache_bytes += bytes;
      g_ctx.bytes -= static_cast<long long>(bytes);
      return;
    }
  }
  dfree(p, bytes);
}

int lfric_vec_create(const lfric_fs *fs, int precision, int nrhs, lfric_vec **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(fs && out && nrhs >= 1, LFRIC_E_ARG, "bad arguments");
  LF_CHECK(precision == LFRIC_R64 || precision == LFRIC_R32, LFRIC_E_UNSUPPORTED, "vector precision");
  auto *v = new lfric_vec;
  v->fs = fs->fs;
  v->precision = precision;
  v->nrhs = nrhs;
  v->n = static_cast<long long>(fs->fs->undf) * nrhs;
  v->bytes = esize(precision) * static_cast<size_t>(v->n);
  if (int rc = vec_buffer_get(&v->d, v->bytes)) {
    delete v;
    return rc;
  }
  LF_HIP(hipMemsetAsync(v->d, 0, v->bytes, cur_stream()));
  *out = v;
  return LFRIC_OK;
}

void lfric_vec_destroy(lfric_vec *v) {
  if (!v) return;
  vec_buffer_put(v->d, v->bytes);
  delete v;
}

static int upload_any(lfric_vec *v, const void *host, int hprec) {
  if (int rc = require_init()) return rc;
  LF_CHECK(v && host, LFRIC_E_ARG, "null argument");
  if (hprec == v->precision) {
    LF_HIP(hipMemcpyAsync(v->d, host, v->bytes, hipMemcpyHostToDevice, cur_stream()));
    LF_HIP(hipStreamSynchronize(cur_stream()));
    return LFRIC_OK;
  }
  void *tmp;
  const size_t hb = esize(hprec) * static_cast<size_t>(v->n);
  if (int rc = ws_get(7, hb, &tmp)) return rc;
  LF_HIP(hipMemcpyAsync(tmp, host, hb, hipMemcpyHostToDevice, cur_stream()));
  if (int rc = launch_convert(v->precision, v->d, hprec, tmp, v->n, 1.0)) return rc;
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}

static int download_any(const lfric_vec *v, void *host, int hprec) {
  if (int rc = require_init()) return rc;
  LF_CHECK(v && host, LFRIC_E_ARG, "null argument");
  if (hprec == v->precision) {
    LF_HIP(hipMemcpyAsync(host, v->d, v->bytes, hipMemcpyDeviceToHost, cur_stream()));
    LF_HIP(hipStreamSynchronize(cur_stream()));
    return LFRIC_OK;
  }
  void *tmp;
  const size_t hb = esize(hprec) * static_cast<size_t>(v->n);
  if (int rc = ws_get(7, hb, &tmp)) return rc;
  if (int rc = launch_convert(hprec, tmp, v->precision, v->d, v->n, 1.0)) return rc;
  LF_HIP(hipMemcpyAsync(host, tmp, hb, hipMemcpyDeviceToHost, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}

int lfric_vec_upload(lfric_vec *v, const double *host) { return upload_any(v, host, LFRIC_R64); }
int lfric_vec_download(const lfric_vec *v, double *host) { return download_any(v, host, LFRIC_R64); }
int lfric_vec_upload_r32(lfric_vec *v, const float *host) { return upload_any(v, host, LFRIC_R32); }
int lfric_vec_download_r32(const lfric_vec *v, float *host) { return download_any(v, host, LFRIC_R32); }

int lfric_vec_size(const lfric_vec *v, int *undf, int *nrhs) {
  LF_CHECK(v, LFRIC_E_ARG, "null vector");
  if (undf) *undf = v->fs->undf;
  if (nrhs) *nrhs = v->nrhs;
  return LFRIC_OK;
}

#define SAME_SHAPE(a, b) LF_CHECK((a) && (b) && (a)->n == (b)->n && (a)->precision == (b)->precision, LFRIC_E_ARG, "vector shapes or precisions differ")

int lfric_vec_set_scalar(lfric_vec *v, double s) {
  if (int rc = require_init()) return rc;
  LF_CHECK(v, LFRIC_E_ARG, "null vector");
  return launch_set(v->precision, v->d, v->n, s);
}
int lfric_vec_axpy(lfric_vec *y, double a, const lfric_vec *x) {
  if (int rc = require_init()) return rc;
  SAME_SHAPE(y, x);
  return launch_axpy(y->precision, y->d, x->d, y->n, a);
}
int lfric_vec_aypx(lfric_vec *y, double a, const lfric_vec *x) {
  if (int rc = require_init()) return rc;
  SAME_SHAPE(y, x);
  return launch_aypx(y->precision, y->d, x->d, y->n, a);
}
int lfric_vec_axpby(lfric_vec *z, double a, const lfric_vec *x, double b, const lfric_vec *y) {
  if (int rc = require_init()) return rc;
  SAME_SHAPE(z, x);
  SAME_SHAPE(z, y);
  return launch_axpby(z->precision, z->d, a, x->d, b, y->d, z->n);
}
int lfric_vec_scale(lfric_vec *v, double a) {
  if (int rc = require_init()) return rc;
  LF_CHECK(v, LFRIC_E_ARG, "null vector");
  return launch_scale(v->precision, v->d, v->n, a);
}
int lfric_vec_copy(lfric_vec *dst, const lfric_vec *src) {
  if (int rc = require_init()) return rc;
  LF_CHECK(dst && src && dst->n == src->n, LFRIC_E_ARG, "vector shapes differ");
  if (dst->precision == src->precision) return launch_copy(dst->precision, dst->d, src->d, dst->n);
  return launch_convert(dst->precision, dst->d, src->precision, src->d, dst->n, 1.0);
}
int lfric_vec_divide(lfric_vec *z, const lfric_vec *x, const lfric_vec *d) {
  if (int rc = require_init()) return rc;
  SAME_SHAPE(z, x);
  SAME_SHAPE(z, d);
  return launch_divide(z->precision, z->d, x->d, d->d, z->n);
}
int lfric_vec_dot(const lfric_vec *x, const lfric_vec *y, double *out) {
  if (int rc = require_init()) return rc;
  SAME_SHAPE(x, y);
  LF_CHECK(out, LFRIC_E_ARG, "null output");
  if (int rc = launch_dot(x->precision, x->d, y->d, x->n, &g_state->red0)) return rc;
  LF_HIP(hipMemcpyAsync(out, &g_state->red0, sizeof(double), hipMemcpyDeviceToHost, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}
int lfric_vec_norm(const lfric_vec *x, double *out) {
  double d = 0;
  if (int rc = lfric_vec_dot(x, x, &d)) return rc;
  *out = std::sqrt(d);
  return LFRIC_OK;
}
int lfric_vec_fill_uniform(lfric_vec *v, unsigned long long seed, unsigned long long stream) {
  if (int rc = require_init()) return rc;
  LF_CHECK(v, LFRIC_E_ARG, "null vector");
  return launch_vec_uniform(v, seed, stream);
}

// ---- operators ------------------------------------------------------------------------------
int lfric_op_create(const lfric_fs *to, const lfric_fs *from, int precision, lfric_op **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(to && from && out, LFRIC_E_ARG, "null argument");
  LF_CHECK(precision == LFRIC_R64 || precision == LFRIC_R32, LFRIC_E_UNSUPPORTED, "operator precision");
  const FunctionSpace &t = *to->fs, &f = *from->fs;
  LF_CHECK(t.ncells == f.ncells && t.nlayers == f.nlayers && t.cell_of_pos == f.cell_of_pos, LFRIC_E_ARG,
           "operator spaces must share one mesh and colouring");
  auto *op = new lfric_op;
  op->to = to->fs;
  op->from = from->fs;
  op->precision = precision;
  // 64 elements per group, each element's entries padded to whole 16-byte chunks
  const size_t V = 16 / esize(precision);
  op->gs = 64 * ((static_cast<size_t>(t.ndf) * f.ndf + V - 1) / V) * V;
  op->ngroups = t.nelem_padded / 64;
  op->row_top = t.top_mask;
  op->bytes = esize(precision) * op->gs * static_cast<size_t>(op->ngroups);
  if (int rc = dmalloc(&op->d, op->bytes)) {
    delete op;
    return rc;
  }
  LF_HIP(hipMemsetAsync(op->d, 0, op->bytes, cur_stream()));
  *out = op;
  return LFRIC_OK;
}

void lfric_op_destroy(lfric_op *op) {
  if (!op) return;
  dfree(op->d, op->bytes);
  dfree(op->rows, op->rows_bytes);
  drop_sym(op);
  delete op;
}

// an operator whose entries change keeps its row-gathered and symmetric-packed copies (those it
// has) in step; new entries that are not symmetric drop the symmetric copy
static int refresh_rows(lfric_op *op, int rc) {
  ++op->version;
  if (rc != LFRIC_OK) return rc;
  if (op->rows)
    if (int r = launch_build_rows(op)) return r;
  if (op->sym) {
    const int r = launch_build_sym(op, op->sym_seg);
    if (r == LFRIC_E_UNSUPPORTED) {
      drop_sym(op);
    } else if (r) {
      return r;
    }
  }
  return LFRIC_OK;
}

static int op_upload_any(lfric_op *op, const void *ls, int prec) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op && ls, LFRIC_E_ARG, "null argument");
  const size_t n = static_cast<size_t>(op->to->ncells) * op->to->nlayers * op->to->ndf * op->from->ndf;
  const size_t esize = prec == LFRIC_R64 ? sizeof(double) : sizeof(float);
  void *tmp = nullptr;
  if (int rc = dmalloc(&tmp, n * esize)) return rc;
  LF_HIP(hipMemcpyAsync(tmp, ls, n * esize, hipMemcpyHostToDevice, cur_stream()));
  int rc = launch_op_from_reference(op, tmp, prec);
  dfree(tmp, n * esize);
  return refresh_rows(op, rc);
}

static int op_download_any(const lfric_op *op, void *ls, int prec) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op && ls, LFRIC_E_ARG, "null argument");
  const size_t n = static_cast<size_t>(op->to->ncells) * op->to->nlayers * op->to->ndf * op->from->ndf;
  const size_t esize = prec == LFRIC_R64 ? sizeof(double) : sizeof(float);
  void *tmp = nullptr;
  if (int rc = dmalloc(&tmp, n * esize)) return rc;
  int rc = launch_op_to_reference(op, tmp, prec);
  if (rc == LFRIC_OK) {
    hipError_t e = hipMemcpy(ls, tmp, n * esize, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(LFRIC_E_HIP, hipGetErrorString(e));
  }
  dfree(tmp, n * esize);
  return rc;
}

int lfric_op_upload(lfric_op *op, const double *ls) { return op_upload_any(op, ls, LFRIC_R64); }
int lfric_op_download(const lfric_op *op, double *ls) { return op_download_any(op, ls, LFRIC_R64); }
int lfric_op_upload_r32(lfric_op *op, const float *ls) { return op_upload_any(op, ls, LFRIC_R32); }
int lfric_op_download_r32(const lfric_op *op, float *ls) { return op_download_any(op, ls, LFRIC_R32); }

int lfric_op_generate_spd(lfric_op *op, unsigned long long seed, int scaled, double shift) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op, LFRIC_E_ARG, "null operator");
  return refresh_rows(op, launch_gen_spd(op, seed, scaled, shift));
}

int lfric_op_generate_uniform(lfric_op *op, unsigned long long seed, unsigned long long stream, double lo,
                              double hi, const double *col_sign) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op, LFRIC_E_ARG, "null operator");
  return refresh_rows(op, launch_gen_uniform(op, seed, stream, lo, hi, col_sign));
}

int lfric_op_convert(const lfric_op *src, lfric_op *dst) {
  if (int rc = require_init()) return rc;
  LF_CHECK(src && dst && src->to == dst->to && src->from == dst->from, LFRIC_E_ARG, "operators differ in shape");
  return refresh_rows(dst, launch_op_convert(src, dst));
}

int lfric_op_assemble_mass_w0(lfric_op *op, const lfric_vec *chi1, const lfric_vec *chi2, const lfric_vec *chi3,
                              int nqp) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op && chi1 && chi2 && chi3, LFRIC_E_ARG, "null argument");
  for (const lfric_vec *c : {chi1, chi2, chi3})
    LF_CHECK(c->precision == LFRIC_R64 && c->nrhs == 1 && c->fs == op->to, LFRIC_E_ARG,
             "assemble_mass_w0: chi must be FP64 fields on the operator's space");
  return refresh_rows(op, launch_assemble_mass_w0(op, static_cast<const double *>(chi1->d),
                                                  static_cast<const double *>(chi2->d),
                                                  static_cast<const double *>(chi3->d), nqp));
}

static int check_chi(const lfric_vec *chi1, const lfric_vec *chi2, const lfric_vec *chi3) {
  LF_CHECK(chi1 && chi2 && chi3, LFRIC_E_ARG, "null argument");
  for (const lfric_vec *c : {chi1, chi2, chi3})
    LF_CHECK(c->precision == LFRIC_R64 && c->nrhs == 1 && c->fs == chi1->fs, LFRIC_E_ARG,
             "assemble: chi must be FP64 fields on one coordinate space");
  return LFRIC_OK;
}

int lfric_op_assemble_mass(lfric_op *op, const lfric_vec *chi1, const lfric_vec *chi2, const lfric_vec *chi3, int nqp,
                           int rehabilitate) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op, LFRIC_E_ARG, "null argument");
  if (int rc = check_chi(chi1, chi2, chi3)) return rc;
  const auto *c1 = static_cast<const double *>(chi1->d), *c2 = static_cast<const double *>(chi2->d),
             *c3 = static_cast<const double *>(chi3->d);
  const FunctionSpace &f = *op->to;
  if (f.ndf == 8) {
    LF_CHECK(chi1->fs == op->to || chi1->fs->dofmap == f.dofmap, LFRIC_E_ARG,
             "assemble: W0 coordinates must share the operator's dofmap");
    return refresh_rows(op, launch_assemble_mass_w0(op, c1, c2, c3, nqp));
  }
  const int kind = f.ndf == 1 ? 0 : f.ndf == 2 ? 3 : 1;  // W3, Wtheta, else W2 (the launcher checks the roles)
  return refresh_rows(op, launch_assemble_fem(op, kind, *chi1->fs, c1, c2, c3, nqp, rehabilitate != 0));
}

int lfric_op_assemble_div(lfric_op *op, const lfric_vec *chi1, const lfric_vec *chi2, const lfric_vec *chi3, int nqp,
                          int rehabilitate) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op, LFRIC_E_ARG, "null argument");
  if (int rc = check_chi(chi1, chi2, chi3)) return rc;
  return refresh_rows(op, launch_assemble_fem(op, 2, *chi1->fs, static_cast<const double *>(chi1->d),
                                              static_cast<const double *>(chi2->d),
                                              static_cast<const double *>(chi3->d), nqp, rehabilitate != 0));
}

int lfric_op_invert_local(const lfric_op *A, lfric_op *Ainv) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && Ainv && A != Ainv, LFRIC_E_ARG, "null or aliased operators");
  return refresh_rows(Ainv, launch_op_invert_local(A, Ainv));
}

static int check_apply(const lfric_op *op, const lfric_vec *x, const lfric_vec *y, bool transpose) {
  LF_CHECK(op && x && y, LFRIC_E_ARG, "null argument");
  const auto &xin = transpose ? op->to : op->from;
  const auto &yout = transpose ? op->from : op->to;
  LF_CHECK(x->fs->undf == xin->undf && y->fs->undf == yout->undf, LFRIC_E_ARG, "vector does not match operator space");
  LF_CHECK(x->precision == op->precision && y->precision == op->precision, LFRIC_E_ARG,
           "vector precision differs from operator precision");
  LF_CHECK(x->nrhs == 1 && y->nrhs == 1, LFRIC_E_ARG, "single right-hand side expected");
  LF_CHECK(x->d != y->d, LFRIC_E_ARG, "x and y must differ");
  return LFRIC_OK;
}

int lfric_lma_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  if (int rc = check_apply(op, x, y, false)) return rc;
  return launch_lma_apply(op, x->d, y->d, nullptr);
}

// Average time of the gather apply's kernels over `reps` applies, with HIP events on the
// library's stream: both phases of a symmetric-packed copy on their own, else the one kernel.
int lfric_op_time_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y, int reps, double *phase_ms) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op && x && y && phase_ms && reps > 0 && reps <= 1000, LFRIC_E_ARG, "bad arguments");
  if (int rc = check_apply(op, x, y, false)) return rc;
  std::vector<hipEvent_t> ev(static_cast<size_t>(reps) * 3, nullptr);
  int rc = LFRIC_OK;
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      rc = fail(LFRIC_E_HIP, "hipEventCreate");
      break;
    }
  bool two = false;
  if (!rc) rc = launch_lma_apply_gather(op, x->d, y->d);  // warm-up (builds the gather lists)
  for (int r = 0; r < reps && !rc; ++r) {
    if (hipEventRecord(ev[3 * r], cur_stream()) != hipSuccess) rc = fail(LFRIC_E_HIP, "event");
    if (!rc) rc = launch_lma_apply_gather_marked(op, x->d, y->d, ev[3 * r + 1], &two);
    if (!rc && hipEventRecord(ev[3 * r + 2], cur_stream()) != hipSuccess) rc = fail(LFRIC_E_HIP, "event");
  }
  double p1 = 0, p2 = 0;
  if (!rc && hipStreamSynchronize(cur_stream()) != hipSuccess) rc = fail(LFRIC_E_HIP, "synchronize");
  for (int r = 0; r < reps && !rc; ++r) {
    float a = 0.f, b = 0.f;
    if (two) {
      if (hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 1]) != hipSuccess ||
          hipEventElapsedTime(&b, ev[3 * r + 1], ev[3 * r + 2]) != hipSuccess)
        rc = fail(LFRIC_E_HIP, "event timing");
    } else if (hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 2]) != hipSuccess) {
      rc = fail(LFRIC_E_HIP, "event timing");
    }
    p1 += a;
    p2 += b;
  }
  for (auto e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc) return rc;
  phase_ms[0] = p1 / reps;
  phase_ms[1] = p2 / reps;
  return LFRIC_OK;
}

// Autotuner of the apply's block order: time the apply (form 0 colour-ordered scatter, 1 gather)
// for each XCD window and keep the fastest in the operator.
int lfric_op_tune(lfric_op *op, int form, int *best_window, double *best_ms) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op && (form == 0 || form == 1), LFRIC_E_ARG, "bad arguments");
  Context &c = ctx();
  const size_t es = esize(op->precision);
  void *x = nullptr, *y = nullptr;
  if (int rc = dmalloc(&x, es * static_cast<size_t>(op->from->undf))) return rc;
  if (int rc = dmalloc(&y, es * static_cast<size_t>(op->to->undf))) {
    dfree(x, es * static_cast<size_t>(op->from->undf));
    return rc;
  }
  lfric_vec xv;
  xv.fs = op->from;
  xv.precision = op->precision;
  xv.n = op->from->undf;
  xv.d = x;
  int rc = launch_vec_uniform(&xv, 1, 3);
  const int windows[] = {0, 1, 2, 4, 8, 16, 32, 64, 256};
  double best = 1e30;
  int bw = op->xwin;
  for (int w : windows) {
    if (rc) break;
    op->xwin = w;
    auto go = [&]() { return form == 1 ? launch_lma_apply_gather(op, x, y) : launch_lma_apply(op, x, y, nullptr); };
    if ((rc = go())) break;
    if (hipEventRecord(c.t0, cur_stream()) != hipSuccess) { rc = fail(LFRIC_E_HIP, "event"); break; }
    for (int r = 0; r < 20 && !rc; ++r) rc = go();
    if (rc) break;
    float ms = 0.f;
    if (hipEventRecord(c.t1, cur_stream()) != hipSuccess || hipEventSynchronize(c.t1) != hipSuccess ||
        hipEventElapsedTime(&ms, c.t0, c.t1) != hipSuccess) {
      rc = fail(LFRIC_E_HIP, "event timing");
      break;
    }
    if (ms / 20 < best) {
      best = ms / 20;
      bw = w;
    }
  }
  op->xwin = bw;
  (void)hipStreamSynchronize(cur_stream());
  dfree(x, es * static_cast<size_t>(op->from->undf));
  dfree(y, es * static_cast<size_t>(op->to->undf));
  if (rc) return rc;
  if (best_window) *best_window = bw;
  if (best_ms) *best_ms = best;
  return LFRIC_OK;
}

int lfric_op_prepare_rows(lfric_op *op) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op, LFRIC_E_ARG, "null argument");
  if (int rc = launch_build_rows(op)) return rc;
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}

int lfric_op_prepare_symmetric(lfric_op *op, int levels_per_thread) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op, LFRIC_E_ARG, "null argument");
  if (int rc = launch_build_sym(op, levels_per_thread)) return rc;
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}

int lfric_op_set_window(lfric_op *op, int window) {
  LF_CHECK(op && window >= 0, LFRIC_E_ARG, "bad arguments");
  op->xwin = window;
  return LFRIC_OK;
}

int lfric_lma_apply_mrhs(const lfric_op *op, const lfric_vec *x, lfric_vec *y, int variant) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op && x && y, LFRIC_E_ARG, "null argument");
  LF_CHECK(x->fs->undf == op->from->undf && y->fs->undf == op->to->undf, LFRIC_E_ARG,
           "vector does not match operator space");
  LF_CHECK(x->precision == op->precision && y->precision == op->precision, LFRIC_E_ARG,
           "vector precision differs from operator precision");
  LF_CHECK(x->nrhs == y->nrhs, LFRIC_E_ARG, "x and y hold different numbers of right-hand sides");
  LF_CHECK(x->d != y->d, LFRIC_E_ARG, "x and y must differ");
  return launch_lma_apply_mrhs(op, x->d, y->d, x->nrhs, variant);
}

int lfric_lma_apply_gather(const lfric_op *op, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  if (int rc = check_apply(op, x, y, false)) return rc;
  return launch_lma_apply_gather(op, x->d, y->d);
}

int lfric_lma_apply_stream(const lfric_op *op, const lfric_vec *x, lfric_vec *y, void *stream) {
  if (int rc = require_init()) return rc;
  const hipStream_t keep = t_stream;
  t_stream = static_cast<hipStream_t>(stream);
  const int rc = lfric_lma_apply(op, x, y);
  t_stream = keep;
  return rc;
}

int lfric_lma_apply_transpose(const lfric_op *op, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  if (int rc = check_apply(op, x, y, true)) return rc;
  return launch_lma_transpose(op, x->d, y->d);
}

int lfric_dg_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  if (int rc = check_apply(op, x, y, false)) return rc;
  return launch_dg_apply(op, x->d, y->d);
}

int lfric_dg_inc_apply(const lfric_op *op, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  if (int rc = check_apply(op, x, y, false)) return rc;
  return launch_dg_inc_apply(op, x->d, y->d);
}

int lfric_extrusion_eta(int kind, int nlayers, double *eta) {
  LF_CHECK(eta, LFRIC_E_ARG, "null argument");
  std::vector<double> e;
  if (int rc = extrusion_eta(kind, nlayers, e)) return rc;
  std::copy(e.begin(), e.end(), eta);
  return LFRIC_OK;
}

int lfric_fs_coordinates(const lfric_fs *fs, int extrusion, double domain_top, double scaled_radius, lfric_vec *chi1,
                         lfric_vec *chi2, lfric_vec *chi3) {
  if (int rc = require_init()) return rc;
  LF_CHECK(fs && chi1 && chi2 && chi3, LFRIC_E_ARG, "null argument");
  for (const lfric_vec *c : {chi1, chi2, chi3})
    LF_CHECK(c->fs == fs->fs && c->precision == LFRIC_R64 && c->nrhs == 1, LFRIC_E_ARG,
             "coordinate fields must be FP64 one-RHS vectors on the space");
  const size_t n = static_cast<size_t>(fs->fs->undf);
  std::vector<double> h(3 * n, 0.0);
  if (int rc = cubedsphere_chi(*fs->fs, extrusion, domain_top, scaled_radius, h.data(), h.data() + n, h.data() + 2 * n))
    return rc;
  LF_HIP(hipMemcpy(chi1->d, h.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  LF_HIP(hipMemcpy(chi2->d, h.data() + n, sizeof(double) * n, hipMemcpyHostToDevice));
  LF_HIP(hipMemcpy(chi3->d, h.data() + 2 * n, sizeof(double) * n, hipMemcpyHostToDevice));
  return LFRIC_OK;
}

int lfric_fs_set_boundary_dofs(lfric_fs *fs, const int *bv) {
  if (int rc = require_init()) return rc;
  LF_CHECK(fs && bv, LFRIC_E_ARG, "null argument");
  FunctionSpace &f = *fs->fs;
  uint32_t bot = 0, top = 0;
  for (int df = 0; df < f.ndf; ++df) {
    LF_CHECK((bv[df] == 0 || bv[df] == 1) && (bv[f.ndf + df] == 0 || bv[f.ndf + df] == 1), LFRIC_E_ARG,
             "boundary_dofs entries must be 0 or 1");
    if (bv[df] == 0) bot |= 1u << df;
    if (bv[f.ndf + df] == 0) top |= 1u << df;
  }
  f.bc_bottom = bot;
  f.bc_top = top;
  return LFRIC_OK;
}

int lfric_fs_get_boundary_dofs(const lfric_fs *fs, int *bv) {
  LF_CHECK(fs && bv, LFRIC_E_ARG, "null argument");
  const FunctionSpace &f = *fs->fs;
  for (int df = 0; df < f.ndf; ++df) {
    bv[df] = (f.bc_bottom >> df) & 1u ? 0 : 1;
    bv[f.ndf + df] = (f.bc_top >> df) & 1u ? 0 : 1;
  }
  return LFRIC_OK;
}

int lfric_enforce_bc(lfric_vec *v) {
  if (int rc = require_init()) return rc;
  LF_CHECK(v, LFRIC_E_ARG, "null argument");
  LF_CHECK(v->nrhs == 1, LFRIC_E_UNSUPPORTED, "enforce_bc takes one-RHS vectors");
  return launch_enforce_bc(*v->fs, v->precision, v->d, nullptr);
}

int lfric_mm_diagonal(const lfric_op *op, lfric_vec *diag) {
  if (int rc = require_init()) return rc;
  LF_CHECK(op && diag && diag->fs->undf == op->to->undf && diag->precision == op->precision, LFRIC_E_ARG,
           "diagonal vector does not match operator");
  return launch_mm_diagonal(op, diag->d);
}

// ---- preconditioners ----------------------------------------------------------------------------
int lfric_prec_create_diagonal(const lfric_vec *diag, lfric_prec **out) {
  LF_CHECK(diag && out, LFRIC_E_ARG, "null argument");
  auto *p = new lfric_prec;
  p->kind = 1;
  p->diag = diag;
  *out = p;
  return LFRIC_OK;
}
int lfric_prec_create_identity(lfric_prec **out) {
  LF_CHECK(out, LFRIC_E_ARG, "null argument");
  *out = new lfric_prec;
  return LFRIC_OK;
}
int lfric_prec_apply(const lfric_prec *p, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  LF_CHECK(p && x && y && x->n == y->n && x->precision == y->precision && x->d != y->d, LFRIC_E_ARG, "bad arguments");
  return prec_apply_any(p, x->precision, x->d, y->d, x->n);
}
void lfric_prec_destroy(lfric_prec *p) { delete p; }

// ---- solvers ------------------------------------------------------------------------------------
void lfric_solver_opts_default(lfric_solver_opts *o) {
  if (!o) return;
  o->r_tol = 1e-8;
  o->a_tol = 1e-16;  // sci_mass_matrix_solver_alg_mod.x90:109
  o->max_iter = 500;
  o->gcrk = 30;
  o->monitor_convergence = 1;
  o->fail_on_non_converged = 0;
  o->check_every = 8;
  o->inner_r_tol = 1e-4;
  o->max_outer = 20;
  o->rho_relax = 1.0;
  o->lmin = 0.124;  // W0 bounds of fs_eigenvalues_mod (w0_lmin, w0_lmax)
  o->lmax = 3.377;
}

static int finish_status(const SolverState &s, const lfric_solver_opts &o, lfric_solver_result *res) {
  int st = LFRIC_OK;
  if (s.status == LFRIC_NAN || s.status == LFRIC_BREAKDOWN)
    st = s.status;
  else if (o.monitor_convergence && !s.converged)
    st = LFRIC_NOT_CONVERGED;
  if (res) {
    res->status = st;
    res->iterations = s.iter;
    res->outer_iterations = 0;
    res->r_nrm_0 = s.r_nrm0;
    res->r_nrm = s.r_nrm;
  }
  if (st == LFRIC_NOT_CONVERGED && !o.fail_on_non_converged) return LFRIC_OK;
  if (st != LFRIC_OK) set_error(st == LFRIC_NAN ? "NaN in solver residual" : st == LFRIC_BREAKDOWN ? "solver breakdown" : "solver did not converge");
  return st;
}

static int prec_diag_ptr(const lfric_prec *P, const lfric_op *A, const void **diag) {
  *diag = nullptr;
  if (!P || P->kind == 0) return LFRIC_OK;
  LF_CHECK(P->kind == 1, LFRIC_E_UNSUPPORTED, "the fused CG takes an identity or diagonal preconditioner");
  LF_CHECK(P->diag->fs->undf == A->to->undf && P->diag->precision == A->precision, LFRIC_E_ARG,
           "preconditioner diagonal does not match the operator");
  *diag = P->diag->d;
  return LFRIC_OK;
}

int lfric_cg_solve(const lfric_op *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                   const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && x && b && opts, LFRIC_E_ARG, "null argument");
  LF_CHECK(A->to == A->from, LFRIC_E_ARG, "CG needs a square operator on one space");
  if (int rc = check_apply(A, x, b, false)) return rc;
  const void *diag;
  if (int rc = prec_diag_ptr(P, A, &diag)) return rc;
  CgArgs a{A, diag, x->d, b->d, x->n, A->precision, false, opts->r_tol, opts->a_tol, opts->max_iter,
           opts->monitor_convergence, opts->check_every};
  SolverState s{};
  if (int rc = cg_device(a, s)) return rc;
  int st = finish_status(s, *opts, res);
  if (res) res->true_rel_res = -1.0;  // not computed here: see lfric_residual_norm
  return st;
}

int lfric_cg_solve_noalloc(const lfric_op *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                           const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && x && b && opts, LFRIC_E_ARG, "null argument");
  LF_CHECK(A->to == A->from, LFRIC_E_ARG, "CG needs a square operator on one space");
  if (int rc = check_apply(A, x, b, false)) return rc;
  const void *diag;
  if (int rc = prec_diag_ptr(P, A, &diag)) return rc;
  unsigned char *base = static_cast<unsigned char *>(g_scratch_ptr[ctx().device]);
  LF_CHECK(base, LFRIC_E_STATE, "cg_solve_noalloc: no scratch address recorded for this device");
  const bool fused = ctx().apply_form == 1;
  const bool listed = A->to->d_gcol != nullptr;
  size_t off[4];
  const size_t need = cg_scratch_layout(esize(A->precision) * static_cast<size_t>(x->n),
                                        fused && listed ? gather_dot_blocks(A) : 1, off);
  LF_CHECK(need <= kScratchBytes, LFRIC_E_UNSUPPORTED,
           "cg_solve_noalloc: the problem needs " + std::to_string(need) + " bytes of scratch, more than its " +
               std::to_string(kScratchBytes) + "-byte maximum (3 vectors of undf values plus the block partials)");
  // the gather form's column lists are built on the first apply into a space (an allocation);
  // this entry point only uses them
  LF_CHECK(!fused || listed, LFRIC_E_STATE,
           "cg_solve_noalloc: the space's gather lists are not built yet (they are built by the first "
           "lfric_lma_apply_gather, lfric_op_prepare_rows or solve on the operator)");
  CgArgs a{A, diag, x->d, b->d, x->n, A->precision, false, opts->r_tol, opts->a_tol, opts->max_iter,
           opts->monitor_convergence, opts->check_every};
  SolverState s{};
  if (int rc = cg_device(a, s, base)) return rc;
  int st = finish_status(s, *opts, res);
  if (res) res->true_rel_res = -1.0;
  return st;
}

int lfric_residual_norm(const lfric_op *A, const lfric_vec *x, const lfric_vec *b, double *nrm) {
  if (int rc = require_init()) return rc;
  if (int rc = check_apply(A, x, b, false)) return rc;
  LF_CHECK(nrm, LFRIC_E_ARG, "null output");
  return true_residual(A, x->d, b->d, x->n, x->precision, nrm);
}

int lfric_ir_cg_solve(const lfric_op *A64, const lfric_op *A32, const lfric_prec *P32, lfric_vec *x,
                      const lfric_vec *b, const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A64 && A32 && x && b && opts, LFRIC_E_ARG, "null argument");
  LF_CHECK(A64->precision == LFRIC_R64 && A32->precision == LFRIC_R32, LFRIC_E_ARG,
           "refinement needs an FP64 and an FP32 operator");
  LF_CHECK(A64->to == A32->to && A64->from == A32->from && A64->to == A64->from, LFRIC_E_ARG,
           "operators must share one square space");
  if (int rc = check_apply(A64, x, b, false)) return rc;
  const void *diag;
  if (int rc = prec_diag_ptr(P32, A32, &diag)) return rc;
  const long long n = x->n;
  void *r64, *b32, *d32;
  if (int rc = ws_get(8, 8 * static_cast<size_t>(n), &r64)) return rc;
  if (int rc = ws_get(5, 4 * static_cast<size_t>(n), &b32)) return rc;
  if (int rc = ws_get(6, 4 * static_cast<size_t>(n), &d32)) return rc;
  // r = b - A x in FP64 and its norm
  auto residual = [&](double *nrm) -> int {
    if (int rc = launch_lma_apply_solver(A64, x->d, r64, nullptr)) return rc;
    if (int rc = launch_cg_residual(LFRIC_R64, r64, b->d, n)) return rc;
    if (int rc = launch_dot(LFRIC_R64, r64, r64, n, &g_state->red0)) return rc;
    LF_HIP(hipMemcpyAsync(g_state_host, g_state, sizeof(SolverState), hipMemcpyDeviceToHost, cur_stream()));
    LF_HIP(hipStreamSynchronize(cur_stream()));
    *nrm = std::sqrt(g_state_host->red0);
    return LFRIC_OK;
  };
  double nr0 = 0, nr = 0;
  if (int rc = residual(&nr0)) return rc;
  nr = nr0;
  int total = 0, outer = 0, status = LFRIC_NOT_CONVERGED;
  if (!(nr0 == nr0)) status = LFRIC_NAN;
  else if (nr0 < 2.220446049250313e-16) status = LFRIC_OK;
  while (status == LFRIC_NOT_CONVERGED && outer < opts->max_outer && total < opts->max_iter) {
    if (nr / nr0 <= opts->r_tol || nr <= opts->a_tol) {
      status = LFRIC_OK;
      break;
    }
    // inner: A32 d = r / ||r|| in FP32 from d = 0
    if (int rc = launch_convert(LFRIC_R32, b32, LFRIC_R64, r64, n, 1.0 / nr)) return rc;
    if (int rc = launch_set(LFRIC_R32, d32, n, 0.0)) return rc;
    const double want = opts->r_tol * nr0 / nr;  // reduction still needed
    CgArgs a{A32, diag, d32, b32, n, LFRIC_R32, true, std::max(opts->inner_r_tol, want * 0.5), 0.0,
             opts->max_iter - total, 1, opts->check_every};
    SolverState s{};
    if (int rc = cg_device(a, s)) return rc;
    total += s.iter;
    ++outer;
    if (s.status == LFRIC_NAN) {
      status = LFRIC_NAN;
      break;
    }
    if (int rc = launch_axpy_mixed(static_cast<double *>(x->d), nr, LFRIC_R32, d32, n)) return rc;
    if (int rc = residual(&nr)) return rc;
    if (!(nr == nr)) status = LFRIC_NAN;
  }
  if (status == LFRIC_NOT_CONVERGED && (nr / nr0 <= opts->r_tol || nr <= opts->a_tol)) status = LFRIC_OK;
  if (res) {
    res->status = status;
    res->iterations = total;
    res->outer_iterations = outer;
    res->r_nrm_0 = nr0;
    res->r_nrm = nr;
    res->true_rel_res = nr0 > 0 ? nr / nr0 : 0.0;
  }
  if (status == LFRIC_NOT_CONVERGED && !opts->fail_on_non_converged) return LFRIC_OK;
  if (status != LFRIC_OK) set_error(status == LFRIC_NAN ? "NaN in refinement residual" : "refinement did not converge");
  return status;
}

}  // extern "C"

// CG on an LMA linear operator (the CG case of lfric_solve); a mass-matrix operator that applies
// boundary conditions takes the unfused iteration with enforce_bc after each apply, as
// mass_matrix_operator_type%apply does (sci_mass_matrix_operator_alg_mod.x90:98-101).
namespace lfric {
int cg_solve_linop(const lfric_linop *L, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                   const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (!L->apply_bc) return lfric_cg_solve(L->A, P, x, b, opts, res);
  const lfric_op *A = L->A;
  LF_CHECK(x && b && opts, LFRIC_E_ARG, "null argument");
  if (int rc = check_apply(A, x, b, false)) return rc;
  const void *diag;
  if (int rc = prec_diag_ptr(P, A, &diag)) return rc;
  CgArgs a{A, diag, x->d, b->d, x->n, A->precision, false, opts->r_tol, opts->a_tol, opts->max_iter,
           opts->monitor_convergence, opts->check_every};
  a.bc = true;
  SolverState s{};
  if (int rc = cg_device(a, s)) return rc;
  int st = finish_status(s, *opts, res);
  if (res) res->true_rel_res = -1.0;
  return st;
}
}  // namespace lfric
