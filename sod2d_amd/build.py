"""Build the sod2d_hip shared library in-tree (sod2d_amd/libsod2d_hip.so) with hipcc
for gfx950, and the Fortran shim with flang when it is available.

Usage: python -m sod2d_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsod2d_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("SODHIP_ARCH", "gfx950")

SOURCES = ["runtime.hip", "ops_element.hip", "ops_lps.hip", "ops_node.hip", "ops_sensor.hip", "ops_stage.hip",
           "ops_mlp.hip", "setup.cpp", "clusters.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result"]
# Packed FP32 VALU forms (v_pk_*_f32) are disabled: all FP32 work is issued as
# scalar-per-lane VALU operations.
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + HIP_FLAGS
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


H5LIB = os.path.join(HERE, "libsod2d_h5.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")


def build_h5() -> str | None:
    """The HDF5 mesh/restart I/O library (host code only, sod2d_amd/libsod2d_h5.so),
    linked against the HDF5 C library under HDF5_DIR.  Returns None when that
    library's header is absent."""
    inc, libdir = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"HDF5 not found under {HDF5_DIR}: libsod2d_h5.so not built", file=sys.stderr)
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = H5LIB + ".tmp"
    cmd = [cxx, "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", f"-I{inc}",
           os.path.join(CSRC, "h5io.cpp"), f"-L{libdir}", "-lhdf5", f"-Wl,-rpath,{libdir}",
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5LIB)
    return H5LIB


def build_fortran() -> str | None:
    """Compile the ISO_C_BINDING shim module and the Fortran host driver."""
    flang = shutil.which("flang") or os.path.join(ROCM, "llvm", "bin", "flang")
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(flang) or not os.path.isdir(fdir):
        return None
    mk = os.path.join(fdir, "Makefile")
    if not os.path.exists(mk):
        return None
    subprocess.run(["make", "-s", "-C", fdir, f"FC={flang}"], check=True)
    return fdir


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-fortran", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.jobs))
    print(build_h5())
    if not a.no_fortran:
        build_fortran()


if __name__ == "__main__":
    sys.exit(main())
