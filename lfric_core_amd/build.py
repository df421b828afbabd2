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


def _compile(src, build_dir, verbose):
    obj = build_dir / (src + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
               "-c", str(CSRC / src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return str(obj)


def build(verbose=False, jobs=None):
    """Compile every source (in parallel, one hipcc per file) and link liblfric_hip.so."""
    from concurrent.futures import ThreadPoolExecutor
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2) - 1))
    with ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(lambda src: _compile(src, build_dir, verbose), SOURCES))
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
FORTRAN_LIB = FORTRAN / "_build" / "liblfric_hip_fortran.a"


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING shim (flang) against the standalone solver-API interfaces into
    a static library; an LFRic build compiles the two shim modules against its own modules
    instead, and the drop-in test (oracle/Makefile.dropin) against the reference's."""
    fc = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
    out = FORTRAN / "_build"
    out.mkdir(exist_ok=True)
    objs = []
    for src in FORTRAN_SOURCES:
        obj = out / (Path(src).stem + ".o")
        cmd = [fc, "-O2", "-fPIC", "-module-dir", str(out), "-I", str(out), "-c", str(FORTRAN / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    if FORTRAN_LIB.exists():
        FORTRAN_LIB.unlink()
    subprocess.run(["ar", "rcs", str(FORTRAN_LIB), *objs], check=True)
    return FORTRAN_LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
    print("built", build_fortran(verbose="-v" in sys.argv))
