"""Builds libtcp_gpu.so in-tree with hipcc for gfx950.

    python -m tcp_allpair_amd.build      # or __graft_entry__.build()

Each .hip file under csrc/ is compiled to an object in _build/ and the objects
are linked into tcp_allpair_amd/libtcp_gpu.so.  Packed FP32 instructions are
disabled for all device code (-packed-fp32-ops).  The host programs under
host/ are built with g++: libtcp_io.so (readers and writers of the reference's
text files) and tcp_run (the standalone driver over libtcp_gpu.so).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libtcp_gpu.so")
ARCH = os.environ.get("TCP_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
SOURCES = ["kernels_pair.hip", "kernels_spme.hip", "kernels_md.hip", "tcp_gpu.hip", "batch.hip",
           "noalloc.hip"]
HOST = os.path.join(HERE, "host")
IO_LIB = os.path.join(HERE, "libtcp_io.so")
RUN = os.path.join(HERE, "tcp_run")
GXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-I", os.path.join(ROOT, "include")]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    _run([GXX, *HOST_FLAGS, "-fPIC", "-shared", os.path.join(HOST, "tcp_io.cpp"), "-o", IO_LIB],
         "libtcp_io.so")
    _run([GXX, *HOST_FLAGS, os.path.join(HOST, "tcp_run.cpp"), os.path.join(HOST, "tcp_io.cpp"),
          "-L", HERE, "-ltcp_gpu", "-Wl,-rpath,$ORIGIN", "-o", RUN], "tcp_run")
    if verbose:
        print("built", LIB, IO_LIB, "and", RUN)
    return LIB


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build of {what} failed:\n{r.stderr}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
