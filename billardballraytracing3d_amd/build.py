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


def _run_all(cmds, verbose):
    """Runs the compile commands side by side; raises on the first that fails."""
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    procs = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for cmd in cmds]
    outs = [p.communicate()[0] for p in procs]
    for cmd, p, out in zip(cmds, procs, outs):
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}")
        elif verbose and out.strip():
            print(out)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    o_host = os.path.join(BUILD, "scene_build.o")
    cmds = [["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, "scene_build.cpp"), "-o", o_host]]
    objs = []
    for name in ("bbrt_hip", "bbrt_mlp", "bbrt_objects", "bbrt_bvh"):
        o = os.path.join(BUILD, name + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, name + ".hip"), "-o", o])
        objs.append(o)
    _run_all(cmds, verbose)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + [o_host], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
