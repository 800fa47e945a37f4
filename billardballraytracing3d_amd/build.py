"""Builds libbbrt_hip.so in-tree for gfx950 (hipcc for the kernels, g++ for host-only code)."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libbbrt_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# FP64 decisions must round exactly as written: no FMA contraction anywhere.
# Packed FP32 instructions are disabled for the device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-pthread"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    elif verbose and r.stdout.strip():
        print(r.stdout)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    o_host = os.path.join(BUILD, "scene_build.o")
    _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, "scene_build.cpp"), "-o", o_host], verbose)
    objs = []
    for name in ("bbrt_hip", "bbrt_mlp", "bbrt_objects"):
        o = os.path.join(BUILD, name + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, name + ".hip"), "-o", o], verbose)
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + [o_host], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
