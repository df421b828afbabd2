"""Build liblfric_hip.so in-tree for gfx950 with hipcc (no JIT cache, no torch extension)."""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "liblfric_hip.so"
SOURCES = ["api.hip", "kernels.hip", "krylov.hip", "solvers.hip", "batch.hip", "cma.hip", "mesh_gen.cpp", "ugrid.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def build(verbose=False):
    objs = []
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    for src in SOURCES:
        obj = build_dir / (src + ".o")
        cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                   "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = str(LIB) + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    return LIB


FORTRAN = PKG / "fortran"
FORTRAN_SOURCES = [
    "standalone/constants_mod.f90", "standalone/log_mod.f90", "standalone/vector_mod.f90",
    "standalone/sci_linear_operator_mod.f90", "standalone/sci_preconditioner_mod.f90",
    "standalone/sci_iterative_solver_mod.f90", "lfric_hip_capi_mod.f90", "lfric_hip_solver_api_mod.f90",
]
FORTRAN_TEST = FORTRAN / "_build" / "drop_in_test"


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING shim with the standalone solver-API modules and the drop-in
    test program (flang), linked against liblfric_hip.so."""
    fc = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
    out = FORTRAN / "_build"
    out.mkdir(exist_ok=True)
    objs = []
    for src in FORTRAN_SOURCES + ["tests/drop_in_test.f90"]:
        obj = out / (Path(src).stem + ".o")
        cmd = [fc, "-O2", "-fPIC", "-module-dir", str(out), "-I", str(out), "-c", str(FORTRAN / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    subprocess.run([fc, "-o", str(FORTRAN_TEST), *objs, "-L", str(PKG), "-llfric_hip",
                    "-Wl,-rpath," + str(PKG), "-Wl,-rpath,$ORIGIN/../.."], check=True)
    return FORTRAN_TEST


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
    print("built", build_fortran(verbose="-v" in sys.argv))
