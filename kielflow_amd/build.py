"""Build libkielflow_amd.so (HIP, gfx950) in-tree.

Usage: python -m kielflow_amd.build   (or kielflow_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkielflow_amd.so")
SOURCES = ["kf_fft.hip", "kf_grid.hip", "kf_poisson.hip", "kf_rhs.hip", "kf_adjoint.hip", "kf_effects.hip",
           "kf_batch.hip", "kf_prof.hip", "kf_capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KF_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # a multiply and an add are fused only within one source expression, so
         # that every instantiation of a pass computes the same numbers
         "-ffp-contract=on",
         # packed FP32 instructions are disabled for every kernel of this library
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    workers = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_io(verbose)
    build_driver(verbose)
    return LIB


IO_SRC = os.path.join(HERE, "host", "kf_io.cpp")
IO_LIB = os.path.join(HERE, "libkielflow_io.so")
HDF5_PREFIX = os.environ.get("KF_HDF5_PREFIX", "/opt/conda")


def build_io(verbose: bool = False):
    """libkielflow_io.so: the reference's HDF5 file layout over the HDF5 C
    library (host code only).  Skipped, with a message, where no HDF5
    headers are installed; the rest of the package does not depend on it."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    h5inc = os.path.join(HDF5_PREFIX, "include")
    h5lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(h5inc, "hdf5.h")):
        print(f"kielflow_amd: no hdf5.h under {h5inc}; libkielflow_io.so not built", file=sys.stderr)
        return None
    tmp = IO_LIB + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-Wextra", "-shared", "-fPIC", "-I", inc,
           "-I", h5inc, IO_SRC, "-L", h5lib, "-lhdf5", f"-Wl,-rpath,{h5lib}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"IO library build failed:\n{r.stderr}")
    os.replace(tmp, IO_LIB)
    if verbose:
        print("built", IO_LIB)
    return IO_LIB


DRIVER_SRC = os.path.join(HERE, "host", "kf_driver.cpp")
DRIVER = os.path.join(HERE, "host", "kf_driver")


def build_driver(verbose: bool = False) -> str:
    """The kielflow-style C++ driver over the reference-interface adaptors
    (host code only: g++ against the C ABIs of the two libraries)."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", inc, DRIVER_SRC, "-L", HERE,
           "-lkielflow_amd", "-lkielflow_io", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stderr}")
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
