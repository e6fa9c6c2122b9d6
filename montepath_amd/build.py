"""Builds libmontepath_hip.so (HIP kernels + C-ABI) for gfx950, in-tree.

Run as ``python -m montepath_amd.build``; ``__graft_entry__.build()`` calls
``build()`` too.  hipcc cross-compiles the device code without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(PKG, "_lib")
LIB_NAME = "libmontepath_hip.so"
SOURCES = ["mp_engine.hip", "mp_api.hip"]
ARCH = os.environ.get("MONTEPATH_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the rocm_hip backend needs ROCm's hipcc to build")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-fvisibility=hidden",
        # no packed FP32 instructions (v_pk_*_f32) anywhere in the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # multiply-adds fuse only within one source expression, so every
        # kernel specialisation computes bit-identical results
        "-ffp-contract=on",
        "-I", INCLUDE,
        "-I", CSRC,
        "-Wall",
        "-Wno-unused-function",
    ]


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, LIB_NAME)
    tmp = out + ".tmp"
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIB_DIR, src.replace(".hip", ".o"))
        cmd = [hipcc(), *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    for obj in objs:
        os.remove(obj)
    build_host(verbose)
    return out


HOST = os.path.join(PKG, "host")
HOST_SELFTEST = os.path.join(LIB_DIR, "backend_selftest")


def build_host(verbose: bool = False) -> str:
    """Compile the C++ host layer (montepath::RocmHipBackend over the C-ABI)
    into backend_selftest, linked against the library built beside it."""
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(hipcc()))), "lib")
    tmp = HOST_SELFTEST + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", INCLUDE, "-I", HOST,
           os.path.join(HOST, "rocm_backend.cpp"), os.path.join(HOST, "backend_selftest.cpp"),
           "-L", LIB_DIR, "-lmontepath_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath-link,{rocm_lib}",
           "-pthread", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, HOST_SELFTEST)
    return HOST_SELFTEST


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
