"""Builds the gfx950 library umwm_amd/lib/libumwm_hip.so (and the Fortran
driver when amdflang is present).

Run as `python -m umwm_amd.build` or call build().  Every HIP source is
compiled for gfx950 with packed-FP32 instructions disabled; host-only C++
is compiled with g++.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libumwm_hip.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FLANG = os.path.join(ROCM, "bin", "amdflang")

HIP_SOURCES = ["ops.hip", "fused.hip", "capi.hip", "mlp.hip", "surrogate.hip"]
CXX_SOURCES = ["setup.cpp", "rawio.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_library(verbose=False, extra_hip_flags=()):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        obj = os.path.join(OBJDIR, s + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, *extra_hip_flags, "-c", os.path.join(CSRC, s), "-o", obj])
    for s in CXX_SOURCES:
        obj = os.path.join(OBJDIR, s + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", obj])
    workers = min(len(jobs), max(1, min(os.cpu_count() or 1, 8)))
    with ThreadPoolExecutor(workers) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [os.path.join(OBJDIR, s + ".o") for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-Wl,-soname,libumwm_hip.so", "-pthread"])
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    """Builds the Fortran driver (umwm_amd/bin/umwm_gpu) against the
    library when amdflang exists."""
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(FLANG) or not os.path.isdir(fdir):
        return None
    srcs = [os.path.join(fdir, s) for s in ("umwm_hip_iface.F90", "umwm_driver.F90")
            if os.path.exists(os.path.join(fdir, s))]
    if len(srcs) < 2:
        return None
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    moddir = os.path.join(OBJDIR, "fmod")
    os.makedirs(moddir, exist_ok=True)
    exe = os.path.join(bindir, "umwm_gpu")
    tmp = exe + ".tmp"
    _run([FLANG, "-O2", "-cpp", "-module-dir", moddir, *srcs, "-o", tmp,
          "-L" + LIBDIR, "-lumwm_hip", "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, exe)
    return exe


def build(verbose=False):
    lib = build_library(verbose=verbose)
    exe = build_fortran(verbose=verbose)
    return lib, exe


if __name__ == "__main__":
    lib, exe = build(verbose="-v" in sys.argv)
    print(lib)
    if exe:
        print(exe)
