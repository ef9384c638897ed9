"""Builds the native parts in-tree.

* melody2d_amd/libmelody2d_hip.so - the HIP library for gfx950 behind
  include/melody2d.h (hipcc, FP64 path compiled with -ffp-contract=off so the
  expression order of the reference is kept)
* melody2d_amd/examples/main_hip, main_facade - the reference's driver over
  the C-ABI and over the C++ façade (g++, linked to the library)
* oracle/liboracle.so              - the CPU checker (test infrastructure)
* oracle/_ref/melody_ref           - the reference solver compiled from its own
  sources, only where /root/reference is present (test infrastructure)
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmelody2d_hip.so")
ORACLE = os.path.join(ROOT, "oracle")
REFERENCE = "/root/reference"

HIP_SOURCES = ["mel_context.hip", "k_step.hip", "k_block.hip", "k_fused.hip", "k_proximity.hip", "k_rigid_lists.hip",
               "k_adaptive.hip", "k_status.hip"]
CXX_SOURCES = ["model_io.cpp", "mel_serve.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [_hipcc()] + HIPCC_FLAGS + ["-I", CSRC, "-I", os.path.join(ROOT, "include")] + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


EXAMPLES = ["main_hip", "main_facade"]


def build_examples(verbose=False):
    ex = os.path.join(HERE, "examples")
    for name in EXAMPLES:
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(ex, name + ".cpp"), "-L", HERE, "-lmelody2d_hip", "-Wl,-rpath,$ORIGIN/..",
               "-o", os.path.join(ex, name)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


def build_oracle(verbose=False):
    subprocess.run(["make", "-s", "-C", ORACLE, "oracle"], check=True)
    if os.path.isdir(os.path.join(REFERENCE, "Code")):
        subprocess.run(["make", "-s", "-C", ORACLE, "ref", "REF=" + REFERENCE], check=True)


def build_all(verbose=False):
    build_library(verbose)
    build_examples(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
