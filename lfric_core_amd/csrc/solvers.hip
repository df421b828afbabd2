// The remaining solvers of LFRic's abstract iterative-solver family, device resident:
// BiCGStab, GCR, Jacobi, Chebyshev and precondition-only
// (components/science/source/solver/sci_iterative_solver_mod.f90), and the solver factory of
// the mass-matrix solve (components/science/source/algorithm/solver/
// sci_mass_matrix_solver_alg_mod.x90:164-181).
//
// Each solver runs as a program of small kernels over a bank of FP64 scalars kept in device
// memory: inner products (FP64-accumulated, two-kernel deterministic sums) write a slot,
// one-thread scalar kernels combine slots with the reference's FP64 formulas, and the vector
// updates read their coefficients from slots.  A device flag stops every kernel that writes
// the solution or the scalars once convergence (or a NaN) is seen, so the host only needs to
// poll it every few iterations and the result equals a per-iteration test.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device_common.h"
#include "kernels.h"

namespace lfric {

namespace {

constexpr int kSolBlocks = 1024;  // fixed reduction grid: sums are the same run to run
constexpr double kEpsRdef = 2.220446049250313e-16;  // epsilon(1.0_r_def), constants_mod

struct ScalarBank {
  double s[32];
  int done, status, converged, iter;
};

// scalar slots
enum : int { RHO, RHO_OLD, ALPHA, OMEGA, BETA, TMP, R0V, TT, TS, RN, SC, ERR, R0, NSLOT };

// ---- kernels --------------------------------------------------------------------------
// partial sums of a.b (and c.d when c is given) per block
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pdot(const T *__restrict__ a, const T *__restrict__ b,
                                                 const T *__restrict__ c, const T *__restrict__ d, long long n,
                                                 double *partials, const ScalarBank *bank) {
  if (bank->done) return;
  double v[2] = {0.0, 0.0};
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    v[0] += static_cast<double>(a[i]) * static_cast<double>(b[i]);
    if (c) v[1] += static_cast<double>(c[i]) * static_cast<double>(d[i]);
  }
  block_partials<2>(v, partials);
}

// bank slots <- block-ordered sums (sqrt for norms)
__global__ __launch_bounds__(kBlock) void k_pfin(const double *__restrict__ partials, int nblocks, ScalarBank *bank,
                                                 int slot0, int slot1, int take_sqrt) {
  if (bank->done) return;
  double s[2] = {0.0, 0.0};
  for (int q = threadIdx.x; q < nblocks; q += kBlock) {
    s[0] += partials[2 * q];
    s[1] += partials[2 * q + 1];
  }
  double tot[2];
  block_sum<2>(s, tot);
  if (threadIdx.x != 0) return;
  bank->s[slot0] = take_sqrt ? sqrt(tot[0]) : tot[0];
  if (slot1 >= 0) bank->s[slot1] = tot[1];
}

enum SOp : int {
  kSet,        // s[d] = imm
  kCopy,       // s[d] = s[a]
  kDiv,        // s[d] = s[a] / s[b]
  kNegMul,     // s[d] = -(s[a] * s[b])
  kRecip,      // s[d] = 1 / s[a]
  kBicgBeta,   // s[d] = (s[a] / s[b]) * (s[c] / s[e])
  kIter,       // iter += 1
  kCheckRel,   // s[d] = s[a] / s[b]; converged when s[d] < imm            (relative test)
  kCheckJac,   // converged when s[a] / s[b] <= imm or s[a] <= s[c]          (jacobi test)
};

// one scalar step of a solver program (FP64, the reference's formula); NaN stops the solve
__global__ void k_sop(ScalarBank *bank, int op, int d, int a, int b, int c, int e, double imm) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || bank->done) return;
  double *s = bank->s;
  switch (op) {
    case kSet: s[d] = imm; break;
    case kCopy: s[d] = s[a]; break;
    case kDiv: s[d] = s[a] / s[b]; break;
    case kNegMul: s[d] = -(s[a] * s[b]); break;
    case kRecip: s[d] = 1.0 / s[a]; break;
    case kBicgBeta: s[d] = (s[a] / s[b]) * (s[c] / s[e]); break;
    case kIter: bank->iter += 1; return;
    case kCheckRel: {
      const double err = s[a] / s[b];
      s[d] = err;
      if (err != err) {
        bank->status = LFRIC_NAN;
        bank->done = 1;
      } else if (err < imm) {
        bank->converged = 1;
        bank->done = 1;
      }
      return;
    }
    case kCheckJac: {
      const double r = s[a];
      if (r != r) {
        bank->status = LFRIC_NAN;
        bank->done = 1;
      } else if (r / s[b] <= imm || r <= s[c]) {
        bank->converged = 1;
        bank->done = 1;
      }
      return;
    }
  }
  if (s[d] != s[d]) {
    bank->status = LFRIC_NAN;
    bank->done = 1;
  }
}

// coefficient: (slot >= 0 ? s[slot] : 1) * k, rounded to the vector precision as the
// r_solver vector types do with their FP64 arguments
struct Coef {
  int slot;
  double k;
};
__device__ inline double coef_value(const ScalarBank *bank, Coef c) { return (c.slot >= 0 ? bank->s[c.slot] : 1.0) * c.k; }

// out = ca * a + cb * b (b may be null); skipped once the solve is done
// (out may be a or b: element i is read and written by one thread only)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_lincomb(T *out, const T *a, Coef ca, const T *b, Coef cb, long long n,
                                                    const ScalarBank *bank) {
  if (bank->done) return;
  const T fa = static_cast<T>(coef_value(bank, ca));
  const T fb = static_cast<T>(b ? coef_value(bank, cb) : 0.0);
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (b)
      out[i] = fa * a[i] + fb * b[i];
    else
      out[i] = fa * a[i];
  }
}

}  // namespace

// ---- host side ----------------------------------------------------------------------------
#define LAUNCHS(kernel, grid, ...) hipLaunchKernelGGL((kernel), dim3(grid), dim3(kBlock), 0, cur_stream(), __VA_ARGS__)

namespace {

struct Ws {
  void *p = nullptr;
  size_t bytes = 0;
};
Ws g_sw[10];
ScalarBank *g_bank = nullptr, *g_bank_host = nullptr;
double *g_spart = nullptr;

int sw_get(int slot, size_t bytes, void **out) {
  Ws &w = g_sw[slot];
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

int bank_init() {
  if (!g_bank) {
    if (int rc = dmalloc(reinterpret_cast<void **>(&g_bank), sizeof(ScalarBank))) return rc;
    if (int rc = dmalloc(reinterpret_cast<void **>(&g_spart), sizeof(double) * 2 * kSolBlocks)) return rc;
    LF_HIP(hipHostMalloc(reinterpret_cast<void **>(&g_bank_host), sizeof(ScalarBank), hipHostMallocDefault));
  }
  ScalarBank z{};
  LF_HIP(hipMemcpyAsync(g_bank, &z, sizeof(ScalarBank), hipMemcpyHostToDevice, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}

int bank_read(ScalarBank &h) {
  LF_HIP(hipMemcpyAsync(g_bank_host, g_bank, sizeof(ScalarBank), hipMemcpyDeviceToHost, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  h = *g_bank_host;
  return LFRIC_OK;
}

// A solver program over vectors of one precision
template <typename T>
struct Prog {
  const lfric_linop *A;
  const lfric_prec *P;
  long long n;
  int prec;
  int vgrid;

  int apply(const T *x, T *y) const { return linop_apply(A, x, y); }
  int precond(const T *x, T *y) const { return prec_apply_any(P, prec, x, y, n); }
  int dot(const T *a, const T *b, int slot) const { return dot2(a, b, nullptr, nullptr, slot, -1); }
  int nrm(const T *a, int slot) const {
    LAUNCHS((k_pdot<T>), kSolBlocks, a, a, static_cast<const T *>(nullptr), static_cast<const T *>(nullptr), n,
            g_spart, g_bank);
    LAUNCHS(k_pfin, 1, g_spart, kSolBlocks, g_bank, slot, -1, 1);
    LF_HIP(hipGetLastError());
    return LFRIC_OK;
  }
  int dot2(const T *a, const T *b, const T *c, const T *d, int s0, int s1) const {
    LAUNCHS((k_pdot<T>), kSolBlocks, a, b, c, d, n, g_spart, g_bank);
    LAUNCHS(k_pfin, 1, g_spart, kSolBlocks, g_bank, s0, s1, 0);
    LF_HIP(hipGetLastError());
    return LFRIC_OK;
  }
  int lin(T *out, const T *a, Coef ca, const T *b = nullptr, Coef cb = {-1, 0.0}) const {
    LAUNCHS((k_lincomb<T>), vgrid, out, a, ca, b, cb, n, g_bank);
    LF_HIP(hipGetLastError());
    return LFRIC_OK;
  }
  int sop(int op, int d, int a = 0, int b = 0, int c = 0, int e = 0, double imm = 0.0) const {
    LAUNCHS(k_sop, 1, g_bank, op, d, a, b, c, e, imm);
    LF_HIP(hipGetLastError());
    return LFRIC_OK;
  }
};

constexpr Coef kOne{-1, 1.0};
constexpr Coef kMinusOne{-1, -1.0};

struct Outcome {
  int status = LFRIC_NOT_CONVERGED;
  int iterations = 0, cycles = 0;
  double r0 = 0.0, rn = 0.0;
};

#define TRY(call)                    \
  do {                               \
    if (int rc_ = (call)) return rc_; \
  } while (0)

// ---- bicgstab_solve (sci_iterative_solver_mod.f90:645-779) ------------------------------------
// Preconditioning acts on the search directions: y = P^-1 t, z = P^-1 s, and x is updated with
// them.  The reference writes "NOT converged" when the last allowed iteration converges
// (:764, iter >= max_iter); here the status follows the convergence test.
template <typename T>
int bicgstab_t(const Prog<T> &g, T *x, const T *b, const lfric_solver_opts &o, Outcome &out) {
  void *w[8];
  for (int i = 0; i < 8; ++i) TRY(sw_get(i, sizeof(T) * static_cast<size_t>(g.n), &w[i]));
  T *r = static_cast<T *>(w[0]), *r0 = static_cast<T *>(w[1]), *p = static_cast<T *>(w[2]),
    *v = static_cast<T *>(w[3]), *t = static_cast<T *>(w[4]), *s = static_cast<T *>(w[5]),
    *y = static_cast<T *>(w[6]), *z = static_cast<T *>(w[7]);
  TRY(bank_init());
  TRY(launch_copy(g.prec, r, b, g.n));                 // r = b
  TRY(g.apply(x, v));                                  // v = A x
  TRY(launch_axpy(g.prec, r, v, g.n, -1.0));           // r = r - v
  TRY(launch_copy(g.prec, r0, r, g.n));
  TRY(g.nrm(r, SC));
  ScalarBank h;
  TRY(bank_read(h));
  out.r0 = out.rn = h.s[SC];
  if (h.s[SC] != h.s[SC]) return out.status = LFRIC_NAN, LFRIC_OK;
  if (h.s[SC] < kEpsRdef) return out.status = LFRIC_OK, LFRIC_OK;
  TRY(g.sop(kSet, SC, 0, 0, 0, 0, std::max(h.s[SC], o.a_tol)));
  TRY(g.sop(kSet, ALPHA, 0, 0, 0, 0, 1.0));
  TRY(g.sop(kSet, OMEGA, 0, 0, 0, 0, 1.0));
  TRY(g.sop(kSet, RHO_OLD, 0, 0, 0, 0, 1.0));
  TRY(launch_set(g.prec, v, g.n, 0.0));
  TRY(launch_set(g.prec, p, g.n, 0.0));
  const int every = std::max(1, o.check_every);
  for (int it = 1; it <= o.max_iter; ++it) {
    TRY(g.sop(kIter, 0));
    TRY(g.dot(r, r0, RHO));                                       // rho = r.r0
    TRY(g.sop(kBicgBeta, BETA, RHO, RHO_OLD, ALPHA, OMEGA));      // beta = (rho/rho_old)(alpha/omega)
    TRY(g.sop(kNegMul, TMP, BETA, OMEGA));
    TRY(g.lin(t, r, kOne, v, Coef{TMP, 1.0}));                    // t = r - beta omega v
    TRY(g.precond(t, y));                                          // y = P^-1 t
    TRY(g.lin(p, p, Coef{BETA, 1.0}, y, kOne));                   // p = beta p + y
    TRY(g.apply(p, v));                                            // v = A p
    TRY(g.dot(r0, v, R0V));
    TRY(g.sop(kDiv, ALPHA, RHO, R0V));                            // alpha = rho / r0.v
    TRY(g.lin(s, r, kOne, v, Coef{ALPHA, -1.0}));                 // s = r - alpha v
    TRY(g.precond(s, z));                                          // z = P^-1 s
    TRY(g.apply(z, t));                                            // t = A z
    TRY(g.dot2(t, t, t, s, TT, TS));                              // tt = t.t, ts = t.s
    TRY(g.sop(kDiv, OMEGA, TS, TT));
    TRY(g.lin(x, x, kOne, z, Coef{OMEGA, 1.0}));                  // x = x + omega z
    TRY(g.lin(x, x, kOne, p, Coef{ALPHA, 1.0}));                  // x = x + alpha p
    TRY(g.lin(r, s, kOne, t, Coef{OMEGA, -1.0}));                 // r = s - omega t
    TRY(g.sop(kCopy, RHO_OLD, RHO));
    if (o.monitor_convergence) {
      TRY(g.nrm(r, RN));
      TRY(g.sop(kCheckRel, ERR, RN, SC, 0, 0, o.r_tol));
    }
    if (it % every == 0 || it == o.max_iter) {
      TRY(bank_read(h));
      if (h.done) break;
    }
  }
  TRY(bank_read(h));
  out.iterations = h.iter;
  out.rn = o.monitor_convergence ? h.s[RN] : out.r0;
  out.status = h.status ? h.status : (h.converged || !o.monitor_convergence) ? LFRIC_OK : LFRIC_NOT_CONVERGED;
  return LFRIC_OK;
}

// ---- gcr_solve (sci_iterative_solver_mod.f90:1288-1408) ---------------------------------------
// Restarted GCR(gcrk) with a right preconditioner: dx starts as P^-1 b, the scale is ||b||,
// the host reads the flag once per restart cycle, and x += dx at the end.
template <typename T>
int gcr_t(const Prog<T> &g, T *x, const T *b, const lfric_solver_opts &o, Outcome &out) {
  const int m = o.gcrk;
  LF_CHECK(m >= 1 && m <= 64, LFRIC_E_ARG, "gcrk must be in [1, 64]");
  void *w[5];
  TRY(sw_get(0, sizeof(T) * static_cast<size_t>(g.n), &w[0]));
  TRY(sw_get(1, sizeof(T) * static_cast<size_t>(g.n), &w[1]));
  TRY(sw_get(2, sizeof(T) * static_cast<size_t>(g.n), &w[2]));
  TRY(sw_get(3, sizeof(T) * static_cast<size_t>(g.n) * m, &w[3]));
  TRY(sw_get(4, sizeof(T) * static_cast<size_t>(g.n) * m, &w[4]));
  T *dx = static_cast<T *>(w[0]), *Ax = static_cast<T *>(w[1]), *res = static_cast<T *>(w[2]);
  T *V = static_cast<T *>(w[3]), *PV = static_cast<T *>(w[4]);
  auto vec = [&](T *base, int j) { return base + static_cast<size_t>(j) * g.n; };
  TRY(bank_init());
  TRY(g.precond(b, dx));                               // dx = P^-1 b
  TRY(g.apply(dx, Ax));
  TRY(launch_copy(g.prec, res, b, g.n));
  TRY(launch_axpy(g.prec, res, Ax, g.n, -1.0));        // res = b - A dx
  TRY(g.nrm(b, SC));                                   // sc_err = ||b||
  ScalarBank h;
  TRY(bank_read(h));
  out.r0 = out.rn = h.s[SC];
  if (h.s[SC] != h.s[SC]) return out.status = LFRIC_NAN, LFRIC_OK;
  if (h.s[SC] < kEpsRdef) return out.status = LFRIC_OK, LFRIC_OK;
  if (o.monitor_convergence) TRY(g.sop(kSet, SC, 0, 0, 0, 0, std::max(h.s[SC], o.a_tol)));
  int cycles = 0;
  for (int it = 1; it <= o.max_iter; ++it) {
    for (int iv = 0; iv < m; ++iv) {
      TRY(g.sop(kIter, 0));
      TRY(g.precond(res, vec(PV, iv)));                            // Pv = P^-1 res
      TRY(g.apply(vec(PV, iv), vec(V, iv)));                       // v = A Pv
      for (int j = 0; j < iv; ++j) {
        TRY(g.dot(vec(V, iv), vec(V, j), ALPHA));                  // alpha = v_iv . v_j
        TRY(g.lin(vec(V, iv), vec(V, iv), kOne, vec(V, j), Coef{ALPHA, -1.0}));
        TRY(g.lin(vec(PV, iv), vec(PV, iv), kOne, vec(PV, j), Coef{ALPHA, -1.0}));
      }
      TRY(g.nrm(vec(V, iv), ALPHA));                               // alpha = ||v||
      TRY(g.sop(kRecip, BETA, ALPHA));                             // beta = 1 / alpha
      TRY(g.lin(vec(V, iv), vec(V, iv), Coef{BETA, 1.0}));
      TRY(g.lin(vec(PV, iv), vec(PV, iv), Coef{BETA, 1.0}));
      TRY(g.dot(res, vec(V, iv), ALPHA));                          // alpha = res . v
      TRY(g.lin(dx, dx, kOne, vec(PV, iv), Coef{ALPHA, 1.0}));     // dx += alpha Pv
      TRY(g.lin(res, res, kOne, vec(V, iv), Coef{ALPHA, -1.0}));   // res -= alpha v
      if (o.monitor_convergence) {
        TRY(g.nrm(res, RN));
        TRY(g.sop(kCheckRel, ERR, RN, SC, 0, 0, o.r_tol));
      }
    }
    cycles = it;
    TRY(bank_read(h));
    if (h.done) break;
  }
  TRY(bank_read(h));
  // x = x + dx (the flag has stopped the solver kernels, so add with a plain kernel)
  TRY(launch_axpy(g.prec, x, dx, g.n, 1.0));
  out.iterations = h.iter;
  out.cycles = cycles;
  out.rn = o.monitor_convergence ? h.s[RN] : out.r0;
  out.status = h.status ? h.status : (h.converged || !o.monitor_convergence) ? LFRIC_OK : LFRIC_NOT_CONVERGED;
  return LFRIC_OK;
}

// ---- jacobi_solve (sci_iterative_solver_mod.f90, jacobi_smod) ---------------------------------
// x = x - rho_relax P^-1 (A x - b); the test (when monitoring) is on ||A x - b|| relative to
// its first value, before the update of that iteration.
template <typename T>
int jacobi_t(const Prog<T> &g, T *x, const T *b, const lfric_solver_opts &o, Outcome &out) {
  void *w[2];
  for (int i = 0; i < 2; ++i) TRY(sw_get(i, sizeof(T) * static_cast<size_t>(g.n), &w[i]));
  T *r = static_cast<T *>(w[0]), *z = static_cast<T *>(w[1]);
  TRY(bank_init());
  TRY(g.sop(kSet, TMP, 0, 0, 0, 0, o.a_tol));
  const int every = std::max(1, o.check_every);
  ScalarBank h;
  for (int it = 1; it <= o.max_iter; ++it) {
    TRY(g.sop(kIter, 0));
    TRY(g.apply(x, r));                                 // r = A x
    TRY(g.lin(r, r, kOne, b, kMinusOne));               // r = r - b
    if (o.monitor_convergence) {
      TRY(g.nrm(r, RN));
      if (it == 1) TRY(g.sop(kCopy, R0, RN));
      TRY(g.sop(kCheckJac, 0, RN, R0, TMP, 0, o.r_tol));
    }
    TRY(g.precond(r, z));                               // z = P^-1 r
    TRY(g.lin(x, x, kOne, z, Coef{-1, -o.rho_relax}));  // x = x - rho z
    if (it % every == 0 || it == o.max_iter) {
      TRY(bank_read(h));
      if (h.done) break;
    }
  }
  TRY(bank_read(h));
  out.iterations = h.converged ? h.iter : o.max_iter;
  out.r0 = h.s[R0];
  out.rn = h.s[RN];
  out.status = h.status ? h.status : (h.converged || !o.monitor_convergence) ? LFRIC_OK : LFRIC_NOT_CONVERGED;
  return LFRIC_OK;
}

// ---- chebyshev_solve (sci_iterative_solver_mod.f90, chebyshev_smod) ----------------------------
// max_iter steps from x = 0 for eigenvalue bounds [lmin, lmax] of P^-1 A.  The weights depend
// on nothing computed on the device, so the host evaluates the reference's recurrence and the
// vector updates take them as constants.  The reference logs the final residual at error level
// whenever fail_on_non_converged is set; here the residual is reported and the status is OK.
template <typename T>
int chebyshev_t(const Prog<T> &g, T *x, const T *b, const lfric_solver_opts &o, Outcome &out) {
  LF_CHECK(o.lmax > o.lmin, LFRIC_E_ARG, "chebyshev needs lmax > lmin");
  void *w[4];
  for (int i = 0; i < 4; ++i) TRY(sw_get(i, sizeof(T) * static_cast<size_t>(g.n), &w[i]));
  T *xp = static_cast<T *>(w[0]), *xo = static_cast<T *>(w[1]), *z = static_cast<T *>(w[2]),
    *r = static_cast<T *>(w[3]);
  TRY(bank_init());
  TRY(launch_set(g.prec, x, g.n, 0.0));
  TRY(launch_set(g.prec, xp, g.n, 0.0));
  TRY(launch_set(g.prec, xo, g.n, 0.0));
  TRY(g.nrm(b, SC));
  const double a1 = 2.0 / (o.lmax - o.lmin), a2 = (o.lmax + o.lmin) / (o.lmax - o.lmin), a_over_b = a1 / a2;
  double wgt = 1.0;
  for (int it = 1; it <= o.max_iter; ++it) {
    TRY(g.apply(xo, z));                                      // z = A xo
    TRY(g.lin(r, b, kOne, z, kMinusOne));                     // r = b - z
    TRY(g.precond(r, z));                                     // z = P^-1 r
    wgt = 1.0 / (1.0 - wgt / (4.0 * a2 * a2));
    const double wab = wgt * a_over_b;
    TRY(g.lin(x, z, Coef{-1, wab}, xo, Coef{-1, wgt}));       // x = wab z + w xo
    TRY(g.lin(x, x, kOne, xp, Coef{-1, 1.0 - wgt}));          // x = x + (1-w) xp
    if (it < o.max_iter) {
      TRY(launch_copy(g.prec, xp, xo, g.n));
      TRY(launch_copy(g.prec, xo, x, g.n));
    }
  }
  TRY(g.apply(x, z));
  TRY(g.lin(z, z, kOne, b, kMinusOne));                       // z = A x - b
  TRY(g.nrm(z, RN));
  ScalarBank h;
  TRY(bank_read(h));
  out.iterations = o.max_iter;
  out.r0 = std::max(1.0, h.s[SC]);
  out.rn = h.s[RN];
  out.status = (out.rn == out.rn) ? LFRIC_OK : LFRIC_NAN;
  return LFRIC_OK;
}

// ---- precondition_only_solve (sci_iterative_solver_mod.f90, precondition_only_smod) ----------
template <typename T>
int prec_only_t(const Prog<T> &g, T *x, const T *b, const lfric_solver_opts &o, Outcome &out) {
  TRY(bank_init());
  TRY(g.precond(b, x));                                 // x = P^-1 b
  out.iterations = 1;
  out.status = LFRIC_OK;
  if (o.monitor_convergence) {
    void *w[2];
    for (int i = 0; i < 2; ++i) TRY(sw_get(i, sizeof(T) * static_cast<size_t>(g.n), &w[i]));
    T *res = static_cast<T *>(w[0]), *Ax = static_cast<T *>(w[1]);
    TRY(g.nrm(b, SC));
    TRY(g.apply(x, Ax));
    TRY(g.lin(res, b, kOne, Ax, kMinusOne));
    TRY(g.nrm(res, RN));
    ScalarBank h;
    TRY(bank_read(h));
    out.r0 = h.s[SC];
    out.rn = h.s[RN];
  }
  return LFRIC_OK;
}

using SolveFn32 = int (*)(const Prog<float> &, float *, const float *, const lfric_solver_opts &, Outcome &);
using SolveFn64 = int (*)(const Prog<double> &, double *, const double *, const lfric_solver_opts &, Outcome &);

int run(int method, const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
        const lfric_solver_opts *o, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && x && b && o, LFRIC_E_ARG, "null argument");
  LF_CHECK(x->precision == A->precision && b->precision == A->precision, LFRIC_E_ARG,
           "vectors must have the operator's precision");
  LF_CHECK(x->n == b->n && x->n == A->space->undf, LFRIC_E_ARG, "vector does not match the operator's space");
  LF_CHECK(x->d != b->d, LFRIC_E_ARG, "x and b must differ");
  Outcome out;
  int rc = LFRIC_OK;
  const long long n = x->n;
  const int grid = grid_for(n, kBlock, 8192);
  if (A->precision == LFRIC_R64) {
    Prog<double> g{A, P, n, LFRIC_R64, grid};
    SolveFn64 fn = method == LFRIC_METHOD_BICGSTAB ? bicgstab_t<double>
                   : method == LFRIC_METHOD_GCR     ? gcr_t<double>
                   : method == LFRIC_METHOD_JACOBI  ? jacobi_t<double>
                   : method == LFRIC_METHOD_CHEBYSHEV ? chebyshev_t<double>
                                                      : prec_only_t<double>;
    rc = fn(g, static_cast<double *>(x->d), static_cast<const double *>(b->d), *o, out);
  } else {
    Prog<float> g{A, P, n, LFRIC_R32, grid};
    SolveFn32 fn = method == LFRIC_METHOD_BICGSTAB ? bicgstab_t<float>
                   : method == LFRIC_METHOD_GCR     ? gcr_t<float>
                   : method == LFRIC_METHOD_JACOBI  ? jacobi_t<float>
                   : method == LFRIC_METHOD_CHEBYSHEV ? chebyshev_t<float>
                                                      : prec_only_t<float>;
    rc = fn(g, static_cast<float *>(x->d), static_cast<const float *>(b->d), *o, out);
  }
  if (rc) return rc;
  if (res) {
    res->status = out.status;
    res->iterations = out.iterations;
    res->outer_iterations = out.cycles;
    res->r_nrm_0 = out.r0;
    res->r_nrm = out.rn;
    res->true_rel_res = out.r0 > 0 ? out.rn / out.r0 : 0.0;
  }
  if (out.status == LFRIC_NOT_CONVERGED && !o->fail_on_non_converged) return LFRIC_OK;
  if (out.status == LFRIC_NAN) set_error("NaN in the solver residual");
  if (out.status == LFRIC_NOT_CONVERGED) set_error("solver did not converge");
  return out.status;
}

}  // namespace
}  // namespace lfric

using namespace lfric;

extern "C" {

int lfric_bicgstab_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                         const lfric_solver_opts *opts, lfric_solver_result *res) {
  return run(LFRIC_METHOD_BICGSTAB, A, P, x, b, opts, res);
}

int lfric_gcr_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                    const lfric_solver_opts *opts, lfric_solver_result *res) {
  return run(LFRIC_METHOD_GCR, A, P, x, b, opts, res);
}

int lfric_jacobi_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                       const lfric_solver_opts *opts, lfric_solver_result *res) {
  return run(LFRIC_METHOD_JACOBI, A, P, x, b, opts, res);
}

int lfric_chebyshev_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                          const lfric_solver_opts *opts, lfric_solver_result *res) {
  return run(LFRIC_METHOD_CHEBYSHEV, A, P, x, b, opts, res);
}

int lfric_prec_only_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                          const lfric_solver_opts *opts, lfric_solver_result *res) {
  return run(LFRIC_METHOD_PREC_ONLY, A, P, x, b, opts, res);
}

}  // extern "C"

extern "C" {

// The solver factory of mass_matrix_solver_alg (sci_mass_matrix_solver_alg_mod.x90:164-181,
// 216-283): one entry point selecting the method, with the solver's own options.
int lfric_solve(int method, const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A, LFRIC_E_ARG, "null argument");
  switch (method) {
    case LFRIC_METHOD_CG:
      LF_CHECK(A->kind == 0, LFRIC_E_UNSUPPORTED, "the fused CG runs on an LMA operator");
      return cg_solve_linop(A, P, x, b, opts, res);
    case LFRIC_METHOD_GMRES: return lfric_gmres_solve(A, P, x, b, opts, res);
    case LFRIC_METHOD_FGMRES: return lfric_fgmres_solve(A, P, x, b, opts, res);
    case LFRIC_METHOD_BICGSTAB:
    case LFRIC_METHOD_GCR:
    case LFRIC_METHOD_PREC_ONLY:
    case LFRIC_METHOD_JACOBI:
    case LFRIC_METHOD_CHEBYSHEV: return run(method, A, P, x, b, opts, res);
  }
  return fail(LFRIC_E_ARG, "unknown solver method");
}

}  // extern "C"
