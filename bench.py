#!/usr/bin/env python3
"""Headline benchmark: solved dofs/s at relative residual 1e-8, mixed precision vs FP64, LMA
GB/s and p50/p99 solve ms on one MI355X.

Workload (one step = one solve): the C96 cubed sphere with 70 quadratically stretched levels,
lowest-order W0 (8 dofs per cell, 3,926,158 dofs), synthetic SPD element matrices
A_e = B_e B_e^T + 0.1 I with B_e ~ N(0,1)/8 and each column scaled by 10^U(-2,2), b ~ U(-1,1),
seed 0, Jacobi-preconditioned CG to ||r||/||r0|| <= 1e-8.  The timed solve is FP32 (r_solver)
CG inside FP64 iterative refinement; the plain FP64 CG solve of the same system is timed
beside it.  The level stretching does not enter the synthetic element matrices.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# thread placement of the OpenMP CPU baseline: one thread per physical core, packed; read by the
# OpenMP runtime when the reference library is first loaded
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the FP32 apply's dominant kernel at C96x70 W0 from the rocprofv3
# FETCH_SIZE / WRITE_SIZE runs recorded in profiles/ (counters scaled by a plain copy of known
# size), per apply form: the colour-ordered scatter (per apply, 4 launches), the gather kernel
# over the row-gathered copy, and the first kernel of the symmetric-packed apply.  Re-measure
# when a kernel changes.
TRAFFIC_BYTES_PER_APPLY_FP32 = {("colour", 96, 70): 4 * (288.4e6 + 16.8e6),
                               ("gather", 96, 70): 1074.05e6 + 15.70e6,
                               ("symmetric", 96, 70): 649.4e6 + 65.9e6}


def apply_bytes(fs, ndf, s_a, s_x):
    """Algorithmic bytes of one LMA apply: matrix once, x once, y once, dofmap once."""
    return fs.ncell_3d * ndf * ndf * s_a + 2 * fs.undf * s_x + 4 * fs.ncells * ndf


def sym_bytes(fs, s_a, s_x):
    """Algorithmic bytes of one symmetric-packed apply: the upper triangle of each element
    (36 of 64 entries) once, x once, y once, dofmap once."""
    return fs.ncell_3d * 36 * s_a + 2 * fs.undf * s_x + 4 * fs.ncells * 8


def sym_phase1_bytes(fs, s):
    """Algorithmic bytes of the first kernel of the symmetric-packed apply (k_sym_units1): every
    element's upper triangle once, x once, the dofmap once, and the 4 outputs of every (cell,
    level) unit written once."""
    return fs.ncell_3d * 36 * s + fs.undf * s + 4 * fs.ncells * 8 + fs.ncells * 4 * (fs.nlayers + 1) * s


def solve_roofline(fs, mres, pres, t_mixed, t_fp64, packed=False):
    """Whole-solve HBM rate: sum over iterations of the PCG iteration's algorithmic bytes plus,
    for the mixed solve, each FP64 refinement step (one FP64 apply + 4 FP64 vector passes),
    divided by the solve time.  Iteration: the apply with p.(A p) (matrix, x, y, dofmap once
    each; the matrix counted as the triangles it reads when the apply is the symmetric-packed
    one), the update (reads x, p, r, A p, diag; writes x, r, z: 8 vector passes) and aypx
    (reads z, p; writes p: 3), i.e. 11 vector passes of undf entries beyond the apply."""
    passes = 11
    ab = (lambda s: sym_bytes(fs, s, s)) if packed else (lambda s: apply_bytes(fs, 8, s, s))
    it32 = ab(4) + passes * fs.undf * 4
    it64 = ab(8) + passes * fs.undf * 8
    ref64 = ab(8) + 4 * fs.undf * 8
    b_mixed = mres.iterations * it32 + (mres.outer_iterations + 1) * ref64
    b_fp64 = pres.iterations * it64
    return {"mixed_gbs": b_mixed / t_mixed / 1e9, "fp64_gbs": b_fp64 / t_fp64 / 1e9,
            "mixed_frac": b_mixed / t_mixed / 1e9 / HBM_PEAK_GBS, "fp64_frac": b_fp64 / t_fp64 / 1e9 / HBM_PEAK_GBS,
            "bytes_per_iteration": {"fp32": it32, "fp64": it64}, "vector_passes_per_iteration": passes,
            "unit": "GB/s"}


def time_applies(L, op, x, y, reps, method="apply"):
    fn = getattr(op, method)
    fn(x, y)
    L.synchronize()
    L.timer_start()
    for _ in range(reps):
        fn(x, y)
    return L.timer_stop() / reps


def order1_report(L, reps, n=48, nl=70):
    """FP32 order-1 W0 multi-RHS applies at C48x70 (3.87M dofs, scaled SPD elements), 1 and 4
    right-hand sides: colour-ordered, gather, and the two phases over the symmetric-packed copy.
    GB/s on each form's own algorithmic bytes (full 729-entry elements, or 378-entry triangles)
    and on full-element bytes."""
    fs = L.FunctionSpace.cubedsphere(n, nl, L.W0, 1)
    op = L.LmaOperator(fs).generate_spd(1, scaled=True).convert(L.R32)
    xs = {nr: L.FieldVector(fs, L.R32, nr).fill_uniform(1, 3) for nr in (1, 4)}
    ys = {nr: L.FieldVector(fs, L.R32, nr) for nr in (1, 4)}

    def timed(nr, variant):
        op.apply_mrhs(xs[nr], ys[nr], variant)
        L.synchronize()
        L.timer_start()
        for _ in range(reps):
            op.apply_mrhs(xs[nr], ys[nr], variant)
        return L.timer_stop() / reps

    out = {}
    for nr in (1, 4):
        out[f"rhs{nr}"] = {"colour_ms": timed(nr, 2), "gather_ms": timed(nr, 1)}
    op.prepare_symmetric()
    out["symmetric_window"] = op.tune(1)[0]  # XCD block order of the first phase, by timing
    for nr in (1, 4):
        r = out[f"rhs{nr}"]
        r["symmetric_ms"] = timed(nr, 1)
        xy = 2 * fs.undf * nr * 4 + fs.ncells * 27 * 4
        full, tri = fs.ncell_3d * 729 * 4 + xy, fs.ncell_3d * 378 * 4 + xy
        r["gbs"] = {"colour": full / (r["colour_ms"] * 1e-3) / 1e9, "gather": full / (r["gather_ms"] * 1e-3) / 1e9,
                    "symmetric": tri / (r["symmetric_ms"] * 1e-3) / 1e9,
                    "symmetric_full_element_equivalent": full / (r["symmetric_ms"] * 1e-3) / 1e9}
    out["symmetric_phases_ms_rhs1"] = op.time_apply(xs[1], ys[1], reps)
    return out


_START_AFFINITY = sorted(os.sched_getaffinity(0))  # before any OpenMP runtime binds this thread


def host_cpu_topology():
    """Physical cores and logical CPUs this process may run on (sysfs topology of the CPUs in
    its affinity mask at start-up, the figures lscpu reports)."""
    cpus = _START_AFFINITY
    cores = set()
    for c in cpus:
        base = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            cores.add((int((base / "physical_package_id").read_text()), int((base / "core_id").read_text())))
        except OSError:
            cores.add((0, c))
    return {"physical_cores": len(cores), "logical_cpus": len(cpus)}


def _ref_cg(R, m, A, diag, b_host, r_tol, max_iter):
    """One FP64 Jacobi-PCG solve from x = 0 with the reference's matrix_vector_code in a
    coloured OpenMP loop (oracle/_ref); returns (seconds, iterations)."""
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    I = ctypes.c_int
    x = np.zeros(m.undf)
    t = time.perf_counter()
    it = R.ref_cg_r64(I(m.ncells), I(m.nlayers), I(4), P(m.ncells_per_colour), P(m.cells_in_colour),
                      I(m.cells_in_colour.shape[1]), I(8), I(m.undf), P(m.dofmap), P(A), P(diag), P(b_host), P(x),
                      ctypes.c_double(r_tol), I(max_iter))
    return time.perf_counter() - t, it


def cpu_baseline(fs, op64, b_host, iters_needed, budget_s, small=None):
    """The reference's own matrix_vector_code (compiled from its Fortran source into
    oracle/_ref/libref.so) in a coloured OpenMP loop, inside cg_solve's operation order, on the
    host cores (OMP_PROC_BIND=close, OMP_PLACES=cores).  Bounded sample of the timed workload: a
    fixed number of CG iterations at full size; the solve time is extrapolated as that
    per-iteration time times the iterations the FP64 GPU solve needed.  `small` (the C12x20
    plain FP64 workload: mesh, matrix, diagonal, b) is solved in full, median of 5 solves after
    one warm-up."""
    lib_path = ROOT / "oracle" / "_ref" / "libref.so"
    if not lib_path.exists():
        return None
    from oracle import oracle as O
    R = ctypes.CDLL(str(lib_path))
    R.ref_cg_r64.restype = ctypes.c_int
    R.ref_num_threads.restype = ctypes.c_int
    m = O.Mesh(int(round(math.sqrt(fs.ncells / 6))), fs.nlayers)
    A3 = op64.download()
    A = O.flat_matrix(A3)
    diag = O.mm_diagonal(m, A3)
    del A3

    t2, _ = _ref_cg(R, m, A, diag, b_host, 1e-30, 2)
    n = max(3, min(iters_needed, int(budget_s / max(t2 / 2, 1e-6))))
    t, _ = _ref_cg(R, m, A, diag, b_host, 1e-30, n)
    per_iter = t / n
    solve_s = per_iter * iters_needed
    out = {
        "value": fs.undf / solve_s,
        "unit": "solved dofs/s",
        "cores": R.ref_num_threads(),
        "kind": "reference",
        "extrapolated": True,
        "host": dict(host_cpu_topology(), omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                     omp_places=os.environ.get("OMP_PLACES")),
        "sample": (f"{n} FP64 Jacobi-PCG iterations of the full {m.undf}-dof system ({t:.1f} s, "
                   f"{per_iter * 1e3:.1f} ms/iteration) with the reference matrix_vector_code in a coloured "
                   f"OpenMP loop; solve time extrapolated = per-iteration time x {iters_needed} iterations "
                   f"of the FP64 solve"),
    }
    if small is not None:
        A3s, bs = small
        ms = O.Mesh(12, 20)
        As, ds = O.flat_matrix(A3s), O.mm_diagonal(ms, A3s)
        _ref_cg(R, ms, As, ds, bs, 1e-8, 500)  # warm-up
        runs = [_ref_cg(R, ms, As, ds, bs, 1e-8, 500) for _ in range(5)]
        med = float(np.median([r[0] for r in runs]))
        out["c12x20_fp64_full_solves"] = {
            "median_ms": med * 1e3, "solved_dofs_per_s": ms.undf / med, "iterations": runs[0][1],
            "undf": ms.undf, "solves": len(runs), "all_ms": [r[0] * 1e3 for r in runs],
            "workload": "C12x20 uniform W0 order 0, A_e = B_e B_e^T + 0.1 I (B_e ~ N(0,1)/8, seed 0), "
                        "b ~ U(-1,1), Jacobi PCG to 1e-8, at most 500 iterations"}
    return out


def small_fp64_report(L):
    """The C12x20 plain FP64 workload (the reference's CPU-runnable case): the GPU FP64 and mixed
    solves (median of 5 after one warm-up), and its element matrices and b on the host for the
    CPU solves beside them."""
    fs = L.FunctionSpace.cubedsphere(12, 20)
    op64 = L.LmaOperator(fs).generate_spd(0)
    op32 = op64.convert(L.R32)
    prec64 = L.DiagonalPreconditioner.from_operator(op64)
    prec32 = L.DiagonalPreconditioner.from_operator(op32)
    b = L.FieldVector(fs).fill_uniform(0, 3)
    x = L.FieldVector(fs)
    out = {}
    for name, solver in (("gpu_fp64", L.ConjugateGradient(op64, prec64, 1e-8, 0.0, 500)),
                         ("gpu_mixed", L.MixedPrecisionCG(op64, op32, prec32, 1e-8, 0.0, 500, inner_r_tol=1e-4))):
        ts = []
        for k in range(6):
            x.set_scalar(0.0)
            L.synchronize()
            t = time.perf_counter()
            res = solver.apply(x, b)
            L.synchronize()
            if k:
                ts.append(time.perf_counter() - t)
        med = float(np.median(ts))
        out[name] = {"median_ms": med * 1e3, "solved_dofs_per_s": fs.undf / med, "iterations": res.iterations}
    return out, (op64.download(), b.export_field())


FP64_VECTOR_PEAK_TFLOPS = 78.6
# floating-point operations per quadrature point of the assembly kernels (Jacobian 144, its
# determinant 14, the point weight 1, then per entry: W0 36 x 4, W2 6 x 3 products J b and 21 x 8,
# W3 2, divergence 6 x 2)
ASSEMBLY_FLOPS_PER_POINT = {"mass_w0": 303, "mass_w2": 345, "mass_w3": 161, "div": 171}


def assembly_report(L, fs, nqp=3):
    """On-device assembly of the lowest-order W0, W2, W3 mass matrices and the weak divergence on
    every cell of the mesh, FP64, nqp^3 Gauss points.  The coordinates are synthetic (uniform
    random W0 fields): the kernels' work does not depend on their values."""
    w2, w3 = fs.sibling(L.W2, 0), fs.sibling(L.W3, 0)
    rng = np.random.default_rng(1)
    chi = [L.FieldVector(fs).import_field(rng.uniform(0, 1, fs.undf)) for _ in range(3)]
    nc3 = fs.ncell_3d
    ops = {"mass_w0": (L.LmaOperator(fs), "assemble_mass", 64),
           "mass_w2": (L.LmaOperator(w2), "assemble_mass", 36),
           "mass_w3": (L.LmaOperator(w3), "assemble_mass", 1),
           "div": (L.LmaOperator(w3, w2), "assemble_div", 6)}
    out = {"cells": nc3, "nqp": nqp, "coordinates": "synthetic uniform random W0 fields"}
    for name, (op, method, nent) in ops.items():
        fn = getattr(op, method)
        fn(*chi, nqp=nqp)
        L.synchronize()
        L.timer_start()
        reps = 3
        for _ in range(reps):
            fn(*chi, nqp=nqp)
        ms = L.timer_stop() / reps
        flops = nc3 * nqp ** 3 * ASSEMBLY_FLOPS_PER_POINT[name]
        wbytes = nc3 * nent * 8 + 3 * fs.undf * 8
        tf = flops / (ms * 1e-3) / 1e12
        out[name] = {"ms": ms, "tflops": tf, "frac_fp64": tf / FP64_VECTOR_PEAK_TFLOPS,
                     "bytes": wbytes, "gbs": wbytes / (ms * 1e-3) / 1e9}
    # the CPU restatement of the same kernels (the reference's assembly kernels need generated
    # configuration modules and are not compiled here), one host core, on a sample of columns
    try:
        from oracle import oracle as O
        ncs = 1024
        dm = fs.dofmap()[:ncs]
        chi_h = [rng.uniform(0, 1, fs.undf) for _ in range(3)]
        cpu = {}
        for name, fn in (("mass_w0", O.mass_matrix_w0), ("mass_w2", O.mass_matrix_w2),
                         ("mass_w3", O.mass_matrix_w3), ("div", O.div_operator)):
            t = time.perf_counter()
            fn(ncs, fs.nlayers, dm, *chi_h, nqp=nqp)
            dt = time.perf_counter() - t
            cpu[name] = {"ms_full_mesh": dt * 1e3 * fs.ncells / ncs}
        out["cpu_port"] = {"kind": "port", "cores": 1, "timings": cpu,
                           "sample": f"{ncs} of {fs.ncells} columns ({fs.nlayers} layers) with the oracle's C restatement, "
                                     f"scaled to the whole mesh"}
    except Exception as e:  # the oracle library is test infrastructure; its absence only drops this field
        out["cpu_port"] = {"error": str(e)}
    return out


def columnwise_report(L, fs, reps):
    """Columnwise (CMA) tridiagonal operator on every W3 column of the mesh (nrow = nlayers,
    bandwidth 3, random diagonally dominant bands): lfric_cma_apply and lfric_cma_apply_inverse,
    FP32 and FP64.  Algorithmic bytes per call: the bands, x and y once each,
    ncells * nrow * (3 + 2) * element size.  Beside them, the reference's own
    columnwise_op_appinv_kernel_code (compiled from source into oracle/_ref) on the same FP32
    data, one host core, on a bounded sample of columns."""
    w3 = fs.sibling(L.W3, 0)
    nc, nl = w3.ncells, w3.nlayers
    rng = np.random.default_rng(0)
    bands = rng.uniform(-1, 1, (nc, nl, 3))
    bands[:, :, 1] += 4.0
    out = {"space": f"W3 columns {nc} x {nl}"}
    for P, tag, esz in ((L.R32, "fp32", 4), (L.R64, "fp64", 8)):
        A = L.ColumnwiseOperator(w3, w3, P, custom=(1, 1, 1, 1)).upload(bands)
        x, y = L.FieldVector(w3, P).fill_uniform(3, 3), L.FieldVector(w3, P)
        alg = nc * nl * 5 * esz
        for name, fn in (("apply", lambda: A.apply(x, y)), ("apply_inverse", lambda: A.apply_inverse(x, y))):
            fn()
            L.synchronize()
            L.timer_start()
            for _ in range(reps):
                fn()
            ms = L.timer_stop() / reps
            gbs = alg / (ms * 1e-3) / 1e9
            out[f"{name}_{tag}"] = {"ms": ms, "alg_bytes": alg, "gbs": gbs, "frac": gbs / HBM_PEAK_GBS}
    lib_path = ROOT / "oracle" / "_ref" / "libref.so"
    if lib_path.exists():
        R = ctypes.CDLL(str(lib_path))
        P_ = lambda a: ctypes.c_void_p(a.ctypes.data)
        ncs = min(nc, 8192)
        b32 = np.ascontiguousarray(bands[:ncs], np.float32)
        p = np.array([1, 1, 1, 1, 3, nl, nl], np.int32)
        dm = np.array([[c * nl + 1] for c in range(ncs)], np.int32)
        ind = np.arange(1, nl + 1, dtype=np.int32)
        xr = rng.uniform(-1, 1, ncs * nl).astype(np.float32)
        yr = np.zeros_like(xr)
        t = time.perf_counter()
        R.ref_cma_appinv_r32(ctypes.c_int(ncs), P_(p), P_(b32), P_(xr), P_(yr), ctypes.c_int(1), ctypes.c_int(ncs * nl),
                             P_(dm), P_(ind))
        dt = time.perf_counter() - t
        out["cpu_reference_apply_inverse_fp32"] = {
            "ms_full_mesh": dt * 1e3 * nc / ncs, "cores": 1, "kind": "reference",
            "sample": f"{ncs} columns of {nl} rows with columnwise_op_appinv_kernel_code ({dt * 1e3:.1f} ms), "
                      f"scaled to {nc} columns"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--edge-cells", type=int, default=96)
    ap.add_argument("--nlayers", type=int, default=70)
    ap.add_argument("--apply-reps", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--apply-only", action="store_true",
                    help="time only the LMA applies and a known-size vector copy (for hardware-counter runs); "
                         "with --workload batched, the batch apply only")
    ap.add_argument("--workload", choices=("single", "batched", "served", "helmholtz"), default="single",
                    help="single: C96x70 mixed vs FP64 solve (default); batched: 64 uneven multi-RHS problems "
                         "in one batch vs sequential FP64 solves; served: Poisson request stream with "
                         "continuous batching vs sequential FP64 solves; helmholtz: only the C48x50 "
                         "Helmholtz-like GMRES record (for kernel-trace and counter runs)")
    ap.add_argument("--no-helmholtz", action="store_true",
                    help="single workload: skip the Helmholtz-like GMRES configuration's record")
    ap.add_argument("--problems", type=int, default=64)
    ap.add_argument("--batched-problems", type=int, default=64,
                    help="single workload: also run the batched configuration with this many problems (0: skip)")
    ap.add_argument("--batch-colour", type=int, default=1, choices=(0, 1),
                    help="batched: the operators' copies (1: gather over the row-gathered copy at lowest order, "
                         "symmetric-packed triangles at order 1) or gather form for every problem (0)")
    ap.add_argument("--apply-form", type=int, default=1, choices=(0, 1),
                    help="LMA apply inside the solvers: 0 colour-ordered scatter, 1 gather with fused p.Ap")
    ap.add_argument("--operator-copy", choices=("symmetric", "rows"), default="symmetric",
                    help="gather form's operator copy: symmetric-packed triangles (two kernels) or the "
                         "row-gathered full element matrices (one kernel)")
    ap.add_argument("--serve-window", type=float, default=60.0, help="seconds of arrivals")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="requests per second")
    ap.add_argument("--serve-limit", type=float, default=30.0, help="wall seconds of admission per server")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the hot path is a "
              "single mesh partition; see the sharding design for more than one GPU)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "served":
        return bench_served(args)
    if args.workload == "helmholtz":
        import lfric_core_amd as L
        L.init(0)
        print(json.dumps({"helmholtz": helmholtz_record(L)}))
        return

    import lfric_core_amd as L
    from lfric_core_amd.solvers import residual_norm

    L.init(0)
    L.set_apply_form(args.apply_form)
    fs = L.FunctionSpace.cubedsphere(args.edge_cells, args.nlayers)
    op64 = L.LmaOperator(fs).generate_spd(0, scaled=True)
    op32 = op64.convert(L.R32)
    prec64 = L.DiagonalPreconditioner.from_operator(op64)
    prec32 = L.DiagonalPreconditioner.from_operator(op32)
    b = L.FieldVector(fs).fill_uniform(0, 3)
    x = L.FieldVector(fs)
    undf = fs.undf

    if args.apply_only:
        xs32, ys32 = L.FieldVector(fs, L.R32).fill_uniform(1, 3), L.FieldVector(fs, L.R32)
        xs64, ys64 = L.FieldVector(fs).fill_uniform(1, 3), L.FieldVector(fs)
        out = {"apply_only": True}
        out["gather_cell_layout"] = {"fp32_ms": time_applies(L, op32, xs32, ys32, args.apply_reps, "apply_gather"),
                                     "fp64_ms": time_applies(L, op64, xs64, ys64, args.apply_reps, "apply_gather")}
        op32.prepare_rows()
        op64.prepare_rows()
        # block-order windows of the apply (the autotuner's search space), FP32 and FP64 gather
        out["windows"] = {}
        for wdw in (0, 1, 4, 16, 64, 256):
            op32.set_window(wdw)
            op64.set_window(wdw)
            out["windows"][wdw] = {"fp32_gather_ms": time_applies(L, op32, xs32, ys32, args.apply_reps, "apply_gather"),
                                   "fp64_gather_ms": time_applies(L, op64, xs64, ys64, args.apply_reps, "apply_gather"),
                                   "fp32_colour_ms": time_applies(L, op32, xs32, ys32, args.apply_reps, "apply")}
        op32.set_window(0)
        op64.set_window(0)
        for method in ("apply", "apply_gather"):
            t32 = time_applies(L, op32, xs32, ys32, args.apply_reps, method)
            t64 = time_applies(L, op64, xs64, ys64, args.apply_reps, method)
            out[method] = {"fp32_ms": t32, "fp64_ms": t64,
                           "fp32_gbs": apply_bytes(fs, 8, 4, 4) / (t32 * 1e-3) / 1e9,
                           "fp64_gbs": apply_bytes(fs, 8, 8, 8) / (t64 * 1e-3) / 1e9}
        # 4 right-hand sides at once (FP32): gather form, colour-ordered FMA, colour-ordered MFMA
        xm, ym = L.FieldVector(fs, L.R32, 4).fill_uniform(5, 3), L.FieldVector(fs, L.R32, 4)
        out["multi_rhs4_fp32"] = {}
        for name, variant in (("gather", 1), ("colour", 2), ("mfma", 3)):
            op32.apply_mrhs(xm, ym, variant)
            L.synchronize()
            L.timer_start()
            for _ in range(args.apply_reps):
                op32.apply_mrhs(xm, ym, variant)
            ms = L.timer_stop() / args.apply_reps
            out["multi_rhs4_fp32"][name] = {"ms": ms, "gbs": apply_bytes(fs, 8, 4, 16) / (ms * 1e-3) / 1e9}
        # symmetric-packed copies (upper triangles; two-phase apply), on operators of their own
        s32, s64 = op64.convert(L.R32), L.LmaOperator(fs).generate_spd(0, scaled=True)
        out["symmetric_packed"] = {}
        for seg in (0, 2, 4, 8):
            s32.prepare_symmetric(seg)
            s64.prepare_symmetric(seg)
            t32 = time_applies(L, s32, xs32, ys32, args.apply_reps, "apply_gather")
            t64 = time_applies(L, s64, xs64, ys64, args.apply_reps, "apply_gather")
            out["symmetric_packed"][seg] = {"fp32_ms": t32, "fp64_ms": t64,
                                            "fp32_tri_gbs": sym_bytes(fs, 4, 4) / (t32 * 1e-3) / 1e9,
                                            "fp64_tri_gbs": sym_bytes(fs, 8, 8) / (t64 * 1e-3) / 1e9}
        del s32, s64
        out["order1_c48x70_fp32"] = order1_report(L, args.apply_reps)
        out["columnwise_w3"] = columnwise_report(L, fs, args.apply_reps)
        out["assembly"] = assembly_report(L, fs)
        for _ in range(args.apply_reps):  # plain copies of a known size (undf * 8 bytes each way)
            ys64.copy(xs64)
        L.synchronize()
        out["copy_bytes_each_way"] = undf * 8
        print(json.dumps(out))
        return

    packed = args.apply_form == 1 and args.operator_copy == "symmetric"
    if packed:  # symmetric-packed copies for the gather apply (set-up, like the operators)
        op32.prepare_symmetric(0)
        op64.prepare_symmetric(0)
    elif args.apply_form == 1:  # row-gathered copies
        op32.prepare_rows()
        op64.prepare_rows()
    # the apply autotuner picks each operator's XCD block-order window (once per operator)
    tuned = {"fp32": op32.tune(args.apply_form), "fp64": op64.tune(args.apply_form)}
    mixed = L.MixedPrecisionCG(op64, op32, prec32, 1e-8, 0.0, 20000, inner_r_tol=1e-4, check_every=16)
    plain = L.ConjugateGradient(op64, prec64, 1e-8, 0.0, 20000, check_every=16)

    def run(solver, k):
        times, res = [], None
        for _ in range(k):
            x.set_scalar(0.0)
            L.synchronize()
            t = time.perf_counter()
            res = solver.apply(x, b)
            L.synchronize()
            times.append(time.perf_counter() - t)
        return times, res

    run(mixed, args.warmup)
    L.synchronize()
    t0 = time.perf_counter()
    mixed_times, mres = run(mixed, args.steps)
    L.synchronize()
    total = time.perf_counter() - t0
    x_mixed = x.export_field()
    mixed_true = residual_norm(op64, x, b) / mres.r_nrm_0

    run(plain, args.warmup)
    plain_times, pres = run(plain, args.steps)
    x_plain = x.export_field()
    plain_true = residual_norm(op64, x, b) / pres.r_nrm_0
    err_mixed_vs_fp64 = float(np.linalg.norm(x_mixed - x_plain) / np.linalg.norm(x_plain))

    # LMA apply bandwidth (HIP events on the library stream)
    xs32, ys32 = L.FieldVector(fs, L.R32).fill_uniform(1, 3), L.FieldVector(fs, L.R32)
    xs64, ys64 = L.FieldVector(fs).fill_uniform(1, 3), L.FieldVector(fs)
    method = "apply_gather" if args.apply_form == 1 else "apply"   # the form the solvers used
    t32 = time_applies(L, op32, xs32, ys32, args.apply_reps, method)
    t64 = time_applies(L, op64, xs64, ys64, args.apply_reps, method)
    b32 = sym_bytes(fs, 4, 4) if packed else apply_bytes(fs, 8, 4, 4)
    b64 = sym_bytes(fs, 8, 8) if packed else apply_bytes(fs, 8, 8, 8)
    gbs32 = b32 / (t32 * 1e-3) / 1e9
    gbs64 = b64 / (t64 * 1e-3) / 1e9
    # per-kernel times of the FP32 apply (HIP events around each kernel, library stream)
    k32 = op32.time_apply(xs32, ys32, args.apply_reps) if args.apply_form == 1 else (t32, 0.0)
    if packed:
        kbytes = sym_phase1_bytes(fs, 4)
        kname = "k_sym_units1<float> (first of the two kernels of the symmetric-packed apply)"
    else:
        kbytes = b32
        kname = ("k_lma_gather_rows_w0<float> (one launch per apply, row-gathered copy)" if args.apply_form == 1
                 else "k_lma_colour<float,8,8,RolesW0K0> (4 colour launches per apply)")
    kgbs = kbytes / (k32[0] * 1e-3) / 1e9

    mean_mixed = total / args.steps
    value = undf / mean_mixed
    line = {
        "metric": "solved dofs/s at rel-res 1e-8 (mixed vs FP64), LMA GB/s, p50/p99 solve ms",
        "value": value,
        "unit": "solved dofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mean_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner CG / f64 refinement",
        "data": "synthetic (counter-based generator, seed 0)",
        "config": {"workload": f"C{args.edge_cells}x{args.nlayers} quadratic-stretched W0 order 0, scaled SPD "
                               f"LMA operator, Jacobi PCG to rel-res 1e-8", "undf": undf,
                   "ncell_3d": fs.ncell_3d, "parallelism": "single mesh partition, one GPU"},
        "fp64_solved_dofs_per_s": undf / (sum(plain_times) / len(plain_times)),
        "mixed_vs_fp64_speedup": (sum(plain_times) / len(plain_times)) / mean_mixed,
        "solve_ms": {"mixed_p50": float(np.percentile(mixed_times, 50) * 1e3),
                     "mixed_p99": float(np.percentile(mixed_times, 99) * 1e3),
                     "fp64_p50": float(np.percentile(plain_times, 50) * 1e3),
                     "fp64_p99": float(np.percentile(plain_times, 99) * 1e3)},
        "iterations": {"mixed_inner": mres.iterations, "mixed_outer": mres.outer_iterations,
                       "fp64": pres.iterations},
        "solve_roofline": solve_roofline(fs, mres, pres, sum(mixed_times) / len(mixed_times),
                                         sum(plain_times) / len(plain_times), packed),
        "verification": {"mixed_true_rel_res_fp64": mixed_true, "fp64_true_rel_res": plain_true,
                         "mixed_vs_fp64_solution_rel_err": err_mixed_vs_fp64},
        "lma_gbs": {"fp32": gbs32, "fp64": gbs64, "fp32_ms": t32, "fp64_ms": t64,
                    "fp32_bytes": b32, "fp64_bytes": b64, "fp32_frac": gbs32 / HBM_PEAK_GBS,
                    "bytes_counted": ("symmetric-packed: upper triangles (36 of 64 entries), x, y, dofmap once"
                                      if packed else "full element matrices, x, y, dofmap once"),
                    "fp32_kernel_ms": {"phase1": k32[0], "phase2": k32[1]},
                    "autotuned_window": {"fp32": tuned["fp32"][0], "fp64": tuned["fp64"][0]}},
        "roofline": {"bound": "hbm", "achieved": kgbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kgbs / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES_PER_APPLY_FP32.get(
                         ("symmetric" if packed else "gather" if args.apply_form == 1 else "colour",
                          args.edge_cells, args.nlayers)),
                     "algorithmic_bytes": kbytes, "kernel_ms": k32[0],
                     "kernel": kname + "; bytes, time and traffic per launch"},
    }
    small, small_host = small_fp64_report(L)
    line["c12x20_fp64"] = small
    if not args.no_cpu_baseline:
        bh = b.export_field()
        line["cpu_baseline"] = cpu_baseline(fs, op64, bh, pres.iterations, args.cpu_budget, small_host)
    del op32, op64, prec32, prec64, xs32, ys32, xs64, ys64, mixed, plain
    if not args.no_helmholtz:  # the Helmholtz-like configuration beside the single solve
        line["helmholtz"] = helmholtz_record(L)
    if args.batched_problems > 0:  # the batched configuration beside the single solve
        say = lambda msg: print(f"[batched] {msg}", file=sys.stderr, flush=True)
        line["batched"] = batched_record(L, args.batched_problems, 1, say)
    print(json.dumps(line))


def _common(args, metric, value, unit, ms, config, dtype):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic (counter-based generator)",
            "config": config}


# divergence sign pattern of the Helmholtz-like configuration: outward normals of the faces
# W, S, E, N, B, T of a W2 cell
HELMHOLTZ_FACE_SIGNS = (-1.0, -1.0, 1.0, 1.0, -1.0, 1.0)


def helmholtz_record(L, reps=3):
    """The Helmholtz-like configuration: C48 cubed sphere, 50 levels, W2 (2,087,424 dofs) and W3
    (691,200 dofs); M2 = C C^T + I, D = face signs x U(0.5, 1.5), M3 SPD, and on W3
    H = M3 + 0.5 D diag(M2)^-1 D^T.  GMRES(30) to 1e-8 with FP32 applies and the BF16
    vertical-column block-Jacobi preconditioner (FP32 arithmetic) inside FP64 refinement with
    FP64 Givens rotations: median of `reps` solves after one warm-up; beside it plain FP64
    GMRES(30) without a preconditioner (median of `reps` solves after one warm-up), and the
    FP32 H apply (HIP events)."""
    w2 = L.FunctionSpace.cubedsphere(48, 50, L.W2, 0)
    w3 = w2.sibling(L.W3, 0)
    M2 = L.LmaOperator(w2).generate_spd(0, shift=1.0)
    D = L.LmaOperator(w3, w2).generate_uniform(0, 5, 0.5, 1.5, HELMHOLTZ_FACE_SIGNS)
    M3 = L.LmaOperator(w3).generate_spd(1, shift=1.0)
    H64 = L.LinearOperator.helmholtz(M3, D, M2.mm_diagonal(L.FieldVector(w2)), 0.5)
    M2s, Ds, M3s = M2.convert(L.R32), D.convert(L.R32), M3.convert(L.R32)
    H32 = L.LinearOperator.helmholtz(M3s, Ds, M2s.mm_diagonal(L.FieldVector(w2, L.R32)), 0.5)
    P = L.ColumnJacobiPreconditioner(H32, L.BF16)
    b = L.FieldVector(w3).fill_uniform(9, 3)
    x = L.FieldVector(w3)

    def solve(solver):
        x.set_scalar(0.0)
        L.synchronize()
        t = time.perf_counter()
        res = solver.apply(x, b)
        L.synchronize()
        return time.perf_counter() - t, res

    mixed = L.MixedPrecisionGMRES(H64, H32, P, 30, 1e-8, 0.0, 50, inner_r_tol=1e-4)
    solve(mixed)
    runs = [solve(mixed) for _ in range(reps)]
    t_mixed = float(np.median([t for t, _ in runs]))
    res = runs[-1][1]
    plain = L.GMRES(H64, L.NullPreconditioner(), 30, 1e-8, 0.0, 200)
    solve(plain)  # warm-up: the FP64 Krylov basis is allocated on the first call
    plain_runs = [solve(plain) for _ in range(reps)]
    t_plain = float(np.median([t for t, _ in plain_runs]))
    res64 = plain_runs[-1][1]
    x32, y32 = L.FieldVector(w3, L.R32).fill_uniform(1, 3), L.FieldVector(w3, L.R32)
    t_apply = time_applies(L, H32, x32, y32, 20, "apply")
    # one H apply: D^T (6 entries per cell-layer) and D read once each, M3 once; x read twice, the
    # W2 intermediate written and read, y written
    nc3 = w3.ncell_3d
    apply_bytes = nc3 * (6 + 6 + 1) * 4 + w3.undf * 3 * 4 + w2.undf * 2 * 4
    return {
        "workload": "C48x50 W2/W3 Helmholtz-like H = M3 + 0.5 D diag(M2)^-1 D^T on W3 (691,200 dofs), GMRES(30) "
                    "to 1e-8: FP32 applies, BF16 column block-Jacobi, FP64 refinement and Givens",
        "mixed_ms": t_mixed * 1e3, "mixed_ms_all": [t * 1e3 for t, _ in runs],
        "solved_dofs_per_s": w3.undf / t_mixed,
        "iterations": {"inner": res.iterations, "outer": res.outer_iterations},
        "true_rel_res": res.true_rel_res,
        "plain_fp64_gmres_no_preconditioner": {"ms": t_plain * 1e3, "iterations": res64.iterations,
                                                "status": res64.status},
        "fp32_h_apply": {"ms": t_apply, "bytes": apply_bytes,
                         "gbs": apply_bytes / (t_apply * 1e-3) / 1e9},
    }


def batch_apply_bytes(probs, copies):
    """Algorithmic bytes of one batch apply: with the operators' copies an order-1 problem reads
    the 378 entries of each element's upper triangle (symmetric-packed copy) and a lowest-order
    one its 64 entries (row-gathered copy); x and y once (nrhs values per dof); the dofmaps once.
    Also the same with every element read in full, and the per-unit outputs that the packed form
    writes and reads back between its two phases (not algorithmic)."""
    def ent(p):
        nd = p["fs"].ndf
        return nd * (nd + 1) // 2 if (copies and nd == 27) else nd * nd

    def xyz(p):
        return 2 * p["fs"].undf * p["nrhs"] * 4 + 4 * p["fs"].ncells * p["fs"].ndf

    alg = sum(p["fs"].ncell_3d * ent(p) * 4 + xyz(p) for p in probs)
    full = sum(p["fs"].ncell_3d * p["fs"].ndf ** 2 * 4 + xyz(p) for p in probs)
    rowsum = sum(2 * p["fs"].ncells * 18 * (p["fs"].nlayers + 1) * p["nrhs"] * 4 for p in probs
                 if copies and p["fs"].ndf == 27)
    return alg, full, rowsum


def batched_apply_only(L, problems, say):
    """The batched configuration's problems in one batch, one round run, then the autotuner's
    applies (block sizes x grid sizes, 6 each): for hardware-counter and kernel-trace runs of
    the batch apply."""
    from lfric_core_amd.batch import Batch, batched_mix, build_problem
    probs = [build_problem(*k) for k in batched_mix(problems)]
    bt = Batch(r_tol=1e-8, a_tol=0.0, max_iter=4000, check_every=8)
    for p in probs:
        bt.add(p["op64"], p["op32"], p["d32"], p["x"], p["b"])
    bt.run(1)
    ms = bt.tune()
    st = bt.stats()
    bt.close()
    alg, full, rowsum = batch_apply_bytes(probs, True)
    say(f"tuned apply {ms:.3f} ms {st}")
    return {"apply_only": True, "problems": len(probs), "batch_apply_ms": ms, "alg_bytes": alg,
            "gbs": alg / (ms * 1e-3) / 1e9, "full_matrix_bytes": full, "row_sum_bytes": rowsum,
            "autotuned": {"block": st["block"], "grid_mult": st["grid_mult"]}}


def batched_record(L, problems, steps, say, colour=True):
    """The batched configuration (64 uneven W0 problems, 4 RHS each) solved as one batch: a
    warm-up batch with the autotuner, then `steps` timed batches (host polls every 8 inner
    iterations), one latency batch polled every 2 iterations (each problem's completion time),
    the batch apply's time and HBM rate, and the same 256 right-hand sides solved one after
    another with plain FP64 CG."""
    from lfric_core_amd.batch import Batch, batched_mix, build_problem
    mix = batched_mix(problems)
    say(f"building {len(mix)} problems")
    probs = [build_problem(*k, fp64_copies=True) for k in mix]
    total_dofs = sum(p["fs"].undf * p["nrhs"] for p in probs)
    say(f"built: {total_dofs} dofs x RHS in total")

    def run_batch(tune, check_every):
        bt = Batch(r_tol=1e-8, a_tol=0.0, max_iter=4000, check_every=check_every)
        bt.config(colour=colour)
        ids = []
        for p in probs:
            p["x"].set_scalar(0.0)
            ids.append(bt.add(p["op64"], p["op32"], p["d32"], p["x"], p["b"]))
        L.synchronize()
        t0 = time.perf_counter()
        done_at, tuned = {}, None
        while True:
            na = bt.run(1)
            if tune and tuned is None:
                tuned = bt.tune()
            now = time.perf_counter() - t0
            for pid in ids:
                if pid not in done_at and bt.finished(pid):
                    done_at[pid] = now
            if na == 0:
                break
        wall = time.perf_counter() - t0
        res = [bt.result(pid) for pid in ids]
        st = bt.stats()
        bt.close()
        return wall, np.array(sorted(done_at.values())), res, st, tuned

    w, _, _, st, tuned = run_batch(True, 8)  # warm-up, and the autotuner's choice
    say(f"warm-up batch {w:.2f} s, autotuned apply {tuned:.3f} ms: {st}")
    times = []
    for _ in range(max(1, steps)):
        wall, _, res, st, _ = run_batch(False, 8)
        times.append(wall)
        say(f"timed batch {wall:.3f} s")
    lat_wall, done, lat_res, _, _ = run_batch(False, 2)
    say(f"latency batch {lat_wall:.3f} s")
    worst = max(r.true_rel_res for rr in res + lat_res for r in rr)
    ok = all(r.status == 0 for rr in res + lat_res for r in rr)
    # the batch's multi-RHS apply of all problems at once (autotuned), and its algorithmic bytes:
    # every element matrix once in FP32, x and y once (nrhs values per dof), the dofmaps once
    bt = Batch(r_tol=1e-8, a_tol=0.0, max_iter=4000, check_every=8)
    bt.config(colour=colour)
    for p in probs:
        p["x"].set_scalar(0.0)
        bt.add(p["op64"], p["op32"], p["d32"], p["x"], p["b"])
    bt.run(1)
    apply_ms = bt.tune()
    bt.close()
    apply_bytes_total, apply_bytes_full, rowsum_bytes = batch_apply_bytes(probs, colour)
    # sequential plain FP64 solves of the same systems, one RHS at a time
    seq_t, seq_it = 0.0, 0
    for p in probs:
        B = p["b"].export_field().reshape(p["fs"].undf, p["nrhs"])
        prec = L.DiagonalPreconditioner.from_operator(p["op64"])
        cg = L.ConjugateGradient(p["op64"], prec, 1e-8, 0.0, 20000, check_every=16)
        for r in range(p["nrhs"]):
            bv = L.FieldVector(p["fs"]).import_field(B[:, r])
            xv = L.FieldVector(p["fs"])
            xv.set_scalar(0.0)
            L.synchronize()
            t0 = time.perf_counter()
            out = cg.apply(xv, bv)
            L.synchronize()
            seq_t += time.perf_counter() - t0
            seq_it += out.iterations
    say(f"sequential FP64 {seq_t:.2f} s")
    wall = float(np.median(times))
    nrhs_total = sum(p["nrhs"] for p in probs)
    return {
        "workload": f"{len(probs)} uneven W0 SPD problems (C12/C24/C48, 10/30/70 levels, order 0/1), element "
                    "matrices B B^T + 0.1 I with B ~ N(0,1)/8, 4 RHS each, Jacobi CG to 1e-8 in FP64 refinement, "
                    "one batch; FP32 and FP64 operators with their row-gathered (order 0) or symmetric-packed "
                    "(order 1) copies",
        "multi_rhs_apply": "operator copies: gather over the row-gathered copy (order 0), symmetric-packed "
                           "upper triangles in two phases (order 1)" if colour else "gather form",
        "problems": len(probs), "rhs_total": nrhs_total, "dofs_total": total_dofs,
        "solved_dofs_per_s": total_dofs / wall, "solves_per_s": len(probs) / wall,
        "rhs_solves_per_s": nrhs_total / wall, "batch_wall_ms": wall * 1e3,
        "batch_wall_ms_all": [t * 1e3 for t in times],
        "latency_ms": {"poll_every_iterations": 2, "batch_wall": lat_wall * 1e3,
                       "problem_done_p50": float(np.percentile(done, 50) * 1e3),
                       "problem_done_p99": float(np.percentile(done, 99) * 1e3),
                       "problem_done_min": float(done.min() * 1e3),
                       "problem_done_p10": float(np.percentile(done, 10) * 1e3)},
        # a problem is done when its slowest RHS is: inner FP32 iterations (all refinement steps)
        # of that RHS per problem, which set the completion times above
        "problem_inner_iterations": {q: float(np.percentile([max(r.iterations for r in rr) for rr in lat_res], v))
                                     for q, v in (("min", 0), ("p10", 10), ("p50", 50), ("max", 100))},
        "sequential_fp64": {"wall_ms": seq_t * 1e3, "dofs_per_s": total_dofs / seq_t,
                            "speedup_of_batch": seq_t / wall, "iterations": seq_it},
        "verification": {"all_converged": ok, "worst_true_rel_res_fp64": worst},
        "autotuned": {"block": st["block"], "grid_mult": st["grid_mult"]},
        "batch_apply": {"ms": apply_ms, "alg_bytes": apply_bytes_total,
                        "gbs": apply_bytes_total / (apply_ms * 1e-3) / 1e9,
                        "frac": apply_bytes_total / (apply_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "full_matrix_bytes": apply_bytes_full,
                        "full_matrix_equivalent_gbs": apply_bytes_full / (apply_ms * 1e-3) / 1e9,
                        "row_sum_bytes": rowsum_bytes},
        "batch_iterations": st["iterations"],
    }


def bench_batched(args):
    import lfric_core_amd as L
    L.init(0)
    say = lambda msg: print(f"[batched] {msg}", file=sys.stderr, flush=True)
    if args.apply_only:
        print(json.dumps(batched_apply_only(L, args.problems, say)))
        return
    rec = batched_record(L, args.problems, args.steps, say, colour=bool(args.batch_colour))
    line = _common(args, "solved dofs/s at rel-res 1e-8 (batched mixed precision vs sequential FP64), p50/p99 "
                         "solve ms", rec["solved_dofs_per_s"], "solved dofs/s", rec["batch_wall_ms"],
                   {"workload": rec["workload"], "multi_rhs_apply": rec["multi_rhs_apply"],
                    "problems": rec["problems"], "rhs_total": rec["rhs_total"], "dofs_total": rec["dofs_total"],
                    "parallelism": "one GPU, one batch"},
                   "f32 inner CG / f64 refinement")
    line.update({k: v for k, v in rec.items() if k not in ("workload", "multi_rhs_apply")})
    print(json.dumps(line))


def bench_served(args):
    import lfric_core_amd as L
    from lfric_core_amd import serving

    L.init(0)
    reqs = serving.poisson_stream(args.serve_rate, args.serve_window, seed=0)
    import ctypes
    free, total = ctypes.c_longlong(), ctypes.c_longlong()
    L.lib().lfric_device_memory(ctypes.byref(free), ctypes.byref(total))
    # the 288 GB of HBM, less what the runtime already holds and 8 GB of head-room
    budget = min(288 * 10**9, free.value) - 8 * 10**9
    srv = serving.SolveServer(budget, starve_s=2.0, check_every=8)
    rep = srv.serve(reqs, wall_limit_s=args.serve_limit)
    srv.close()
    done = [r for r in reqs if r.status == 0]
    worst = max((r.true_rel_res / r.tol for r in done), default=float("nan"))
    # the same stream served one request at a time with plain FP64 CG per RHS, in arrival order
    seq = serving.poisson_stream(args.serve_rate, args.serve_window, seed=0)
    t0 = time.perf_counter()
    seq_lat, seq_done, seq_rhs = [], 0, 0
    for r in seq:
        now = time.perf_counter() - t0
        if now >= args.serve_limit:
            break
        if r.t_arrive > now:
            time.sleep(r.t_arrive - now)
        fs = L.FunctionSpace.cubedsphere(r.n, r.nl, L.W0, r.order)
        op64 = L.LmaOperator(fs).generate_spd(r.seed, scaled=True)
        prec = L.DiagonalPreconditioner.from_operator(op64)
        bm = L.FieldVector(fs, L.R64, serving.padded_rhs(r.nrhs)).fill_uniform(r.seed, 3)
        B = bm.export_field().reshape(fs.undf, -1)
        cg = L.ConjugateGradient(op64, prec, r.tol, 0.0, 20000, check_every=16)
        for k in range(r.nrhs):
            xv = L.FieldVector(fs)
            xv.set_scalar(0.0)
            cg.apply(xv, L.FieldVector(fs).import_field(B[:, k]))
        L.synchronize()
        seq_lat.append((time.perf_counter() - t0 - r.t_arrive) * 1e3)
        seq_done += 1
        seq_rhs += r.nrhs
        del op64, prec, bm
    seq_wall = time.perf_counter() - t0
    line = _common(args, "served solves/s with p50/p99 latency (continuous batching, mixed precision, vs "
                         "sequential plain FP64)", rep.solves_per_s, "solves/s", rep.wall_s * 1e3,
                   {"workload": f"Poisson {args.serve_rate:g} req/s for {args.serve_window:g} s (seed 0), C24/C48/C96/"
                                "C192 (0.4/0.3/0.2/0.1), 30/70 levels, order 0/1, 1-8 RHS, tol 1e-6/1e-8, scaled "
                                "SPD elements, 288 GB HBM budget",
                    "requests": len(reqs), "admission_wall_s": args.serve_limit,
                    "parallelism": "one GPU, continuous batching"},
                   "f32 inner CG / f64 refinement")
    line.update({
        "completed": rep.completed, "unfinished": rep.unfinished, "rejected_too_large": rep.rejected,
        "rhs_per_s": rep.rhs_per_s, "dofs_per_s": rep.dofs_per_s,
        "latency_ms": {"p50": rep.p50_ms, "p99": rep.p99_ms, "queue_p50": rep.queue_p50_ms,
                       "queue_p99": rep.queue_p99_ms},
        "peak_admitted_bytes": rep.peak_bytes,
        "verification": {"worst_true_rel_res_over_tol": worst},
        "sequential_fp64": {"completed": seq_done, "solves_per_s": seq_done / seq_wall, "rhs_per_s": seq_rhs / seq_wall,
                            "latency_ms_p50": float(np.percentile(seq_lat, 50)) if seq_lat else None,
                            "latency_ms_p99": float(np.percentile(seq_lat, 99)) if seq_lat else None},
        "admission": "arrival order with backfill; a request waiting over 2 s blocks later admissions; requests "
                     "larger than the budget are rejected; admission stops at the wall limit and admitted "
                     "requests drain",
    })
    print(json.dumps(line))


if __name__ == "__main__":
    main()
