"""Build libfd2.so (HIP kernels + C host library) and libfd2_fortran.so (the
gfortran-ABI entry points) in-tree for gfx950.

    python -m hybridmantis_amd.build          # or hybridmantis_amd.build.build()

hipcc compiles the kernels and the host library into one shared object next
to this file; nothing is installed elsewhere.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfd2.so")

HIP_SOURCES = ["fd2_kernels.hip", "fd2_host.cpp", "fd2_mlp.hip"]
CXX_SOURCES = ["fd2_io.cpp"]
FORTRAN_SOURCE = "fd2_fortran.cpp"
FORTRAN_LIB = os.path.join(HERE, "libfd2_fortran.so")

ARCH = os.environ.get("FD2_OFFLOAD_ARCH", "gfx950")

# Every HIP translation unit is built with the packed-FP32 target feature
# disabled (no v_pk_{add,mul,fma}_f32).
# -ffp-contract=off: no implicit FMA, so the transport's float expressions
# round exactly like the CPU restatement of the same expressions.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
             "-I", os.path.join(ROOT, "include")]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libfd2")


def build(verbose: bool = False, jobs: int = 4, out: str | None = None,
          defines: tuple = ()) -> str:
    """Compile and link libfd2.so (or `out`, with extra -D `defines`)."""
    hipcc = _hipcc()
    lib = out or LIB
    objdir = os.path.join(HERE, "_build" if out is None else "_build_" +
                          os.path.splitext(os.path.basename(out))[0])
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES + CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        if src in HIP_SOURCES:
            cmd = [hipcc, "-x", "hip", *HIP_FLAGS, *[f"-D{d}" for d in defines],
                   "-c", path, "-o", obj]
        else:  # host-only C++ (file formats)
            cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-fPIC",
                   "-Wall", "-I", os.path.join(ROOT, "include"), "-c", path,
                   "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        # each job's output is collected whole, so no pipe can fill and stall it
        return subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)

    with ThreadPoolExecutor(max(1, jobs)) as pool:
        results = list(pool.map(compile_one, cmds))
    failed = []
    for cmd, r in zip(cmds, results):
        if r.stdout and (verbose or r.returncode):
            sys.stderr.write(r.stdout)
        if r.returncode:
            failed.append(cmd[-3])
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = lib + ".tmp"
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, lib)
    if out is None:
        build_fortran(objdir, verbose)
    return lib


def build_fortran(objdir: str, verbose: bool = False) -> str:
    """libfd2_fortran.so: the gfortran-ABI entry points, linked against the
    libfd2.so beside it."""
    gxx = shutil.which("g++") or "g++"
    obj = os.path.join(objdir, "fd2_fortran.o")
    cmds = [[gxx, "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"),
             "-c", os.path.join(CSRC, FORTRAN_SOURCE), "-o", obj],
            [gxx, "-shared", "-fPIC", "-o", FORTRAN_LIB + ".tmp", obj, "-L", HERE, "-lfd2",
             "-Wl,-rpath,$ORIGIN"]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(FORTRAN_LIB + ".tmp", FORTRAN_LIB)
    return FORTRAN_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
