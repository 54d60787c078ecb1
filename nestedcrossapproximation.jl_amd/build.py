"""Builds the in-tree native libraries of the package.

  lib/libnca_hip.so   HIP kernels for gfx950 + the C-ABI of include/nca_hip.h
  lib/libnca_tree.so  host-side trees, partition and Phase-1 pivots (g++, OpenMP)

Run as a script or call build(). Every call compiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are kept out of every kernel.
HIP_COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
              "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
# Translation units and their extra flags. nca_compress must not fuse products:
# its pivots are compared bit for bit with the oracle.
HIP_UNITS = [
    ("nca_api.hip", []),
    ("nca_compress.hip", ["-ffp-contract=off"]),
    ("nca_matvec.hip", []),
    ("nca_partition.hip", ["-ffp-contract=off"]),
    ("nca_tree.hip", ["-ffp-contract=off"]),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_tree():
    os.makedirs(LIB, exist_ok=True)
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o",
          os.path.join(LIB, "libnca_tree.so"), os.path.join(HOST, "nca_tree.cpp")])


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in HIP_UNITS:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", os.path.join(LIB, "libnca_hip.so"), *objs])


def build():
    build_tree()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "tree" in what:
        build_tree()
    elif "hip" in what:
        build_hip()
    else:
        build()
