"""The Fortran drop-in layer on the GPU, driven by LFRic's own solver module.

tests/fortran/drop_in_test.f90 is built by oracle/Makefile.dropin against the reference's
sci_iterative_solver_mod.f90 (and its vector_mod, sci_preconditioner_mod, log_mod and constants_mod)
compiled from the LFRic sources.  LFRic's conjugate_gradient_type, BiCGstab_type and gmres_type then
run on the HIP vector, operator and preconditioner types, in real64 and in r_solver = real32, and
must take the oracle's iteration counts; the device-resident solves must agree with them, and the
columnwise-operator bindings work from Fortran."""
import json
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_fortran_drop_in(orc):
    exe = ROOT / "oracle" / "_ref" / "dropin" / "drop_in_test"
    assert exe.exists(), "Fortran drop-in test not built (oracle/Makefile.dropin)"
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    r = {}
    for line in out.stdout.splitlines():
        if line.startswith("{"):
            r.update(json.loads(line))
    # LFRic's log_mod lines, e.g. "...:INFO : cg converged after     57 iterations"
    cg_logged = [int(v) for v in re.findall(r"cg converged after\s+(\d+) iterations", out.stdout)]
    bs_logged = [int(v) for v in re.findall(r"bicgstab:converged in\s+(\d+) iters", out.stdout)]

    m = orc.Mesh(12, 20)
    A = orc.gen_spd(8, m.ncell_3d, 0)
    b = orc.gen_uniform(m.undf, 0, 3)
    diag = orc.mm_diagonal(m, A)
    _, reso, _ = orc.cg(m, A, b, diag=diag)
    assert r["undf"] == m.undf
    # real64: LFRic's CG made one apply for r0 and one per iteration
    assert cg_logged[0] == r["stock_cg_applies"] - 1 == r["hip_cg_iterations"] == reso.iterations
    assert r["stock_vs_hip_cg_rel_diff"] <= 1e-10
    assert r["stock_gmres_vs_hip_cg_rel_diff"] <= 1e-7
    xb, resb, _ = orc.solve_callbacks("bicgstab", m.undf, lambda v: orc.lma_apply(m, A, v),
                                      lambda v: v / diag, b, r_tol=1e-10, max_iter=300)
    # BiCGStab: one apply for r0 and two per iteration
    assert bs_logged == [resb.iterations]
    assert (r["stock_bicgstab_applies"] - 1) // 2 == r["hip_bicgstab_iterations"] == resb.iterations
    assert r["stock_vs_hip_bicgstab_rel_diff"] <= 1e-9
    # copy() inside LFRic's BiCGStab leaves no device vector behind
    assert r["device_bytes_before_bicgstab"] == r["device_bytes_after_bicgstab"]

    # r_solver: LFRic's CG on real32 vectors over the real32 upload of the same operator
    A32 = A.astype(np.float32).astype(np.float64)
    diag32 = orc.mm_diagonal(m, A32).astype(np.float32)
    x32, it32, ok = orc.cg_r32(m, A32, b.astype(np.float32), diag32, r_tol=1e-5, a_tol=0.0)
    assert ok
    assert cg_logged[1] == r["r_solver_cg_applies"] - 1 == it32
    assert r["r_solver_vs_hip_cg_rel_diff"] <= 1e-4
    assert abs(r["r_solver_x_norm"] - np.linalg.norm(x32.astype(np.float64))) <= 1e-5 * r["r_solver_x_norm"]

    # the columnwise bindings from Fortran: a W3 operator's tridiagonal solve undoes its apply,
    # directly and through the preconditioner wrapper
    assert r["cma_solve_undoes_apply_rel_diff"] <= 1e-13
    assert r["cma_preconditioner_rel_diff"] <= 1e-13
