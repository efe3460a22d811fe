"""Builds libsodabayes.so (gfx950) in-tree with hipcc.

Usage: python -m sodabayes_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsodabayes.so")
SOURCES = ["sb_api.hip", "sb_setup.hip", "sb_sweep.hip", "sb_batch.hip", "sb_predict.hip", "sb_single.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # no FMA contraction: the reference's arithmetic is reproduced operation by
    # operation; kernels that want FMA call fma() explicitly
    "-ffp-contract=off",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", LIB + ".tmp", *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libsodabayes.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


FORTRAN = os.environ.get("FORTRAN", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "build")


def build_fortran() -> str | None:
    """Compiles the iso_c_binding module and its host-side check program
    (fortran/build/binding_check) against libsodabayes.so."""
    if not os.path.exists(FORTRAN):
        return None
    os.makedirs(FBUILD, exist_ok=True)
    exe = os.path.join(FBUILD, "binding_check")
    cmd = [FORTRAN, "-O1", "-J", FBUILD, "-o", exe, os.path.join(FDIR, "soda_gpu.f90"),
           os.path.join(FDIR, "binding_check.f90"), "-L", HERE, "-lsodabayes", "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("Fortran binding failed to build")
    return exe


REF_SRC = os.environ.get("SODABAYES_REF_SRC", "/root/reference/src")
# the reference's host modules the program keeps (its SODA module soda.f90 is
# replaced by soda_gpu_host and not linked); dependency order
REF_HOST_MODULES = ["globals", "parallelrng", "utilities", "useroptions", "models"]


def build_fortran_host() -> str | None:
    """Builds fortran/build/sodabayes_gpu: the reference program with its
    SODA branch on the GPU.  The reference's host modules (option parsing,
    readers, writers, the -SODAOff sampler) are compiled from its sources
    where they lie (REF_SRC); the binding, the glue module and the main
    program are this package's.  Returns None where the reference sources
    or the Fortran compiler are absent."""
    if not (os.path.exists(FORTRAN) and os.path.isdir(REF_SRC)):
        return None
    hdir = os.path.join(FBUILD, "host")
    os.makedirs(hdir, exist_ok=True)
    srcs = [os.path.join(REF_SRC, m + ".f90") for m in REF_HOST_MODULES]
    srcs += [os.path.join(FDIR, f) for f in ("soda_gpu.f90", "soda_gpu_host.f90", "sodabayes_gpu.f90")]
    objs = []
    for src in srcs:
        obj = os.path.join(hdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        r = subprocess.run([FORTRAN, "-O2", "-fopenmp", "-J", hdir, "-c", src, "-o", obj],
                           capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"Fortran host failed to compile {os.path.basename(src)}")
        objs.append(obj)
    exe = os.path.join(FBUILD, "sodabayes_gpu")
    r = subprocess.run([FORTRAN, "-fopenmp", "-o", exe, *objs, "-L", HERE, "-lsodabayes",
                        "-Wl,-rpath,$ORIGIN/../.."], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("Fortran host failed to link")
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran())
    print(build_fortran_host())
