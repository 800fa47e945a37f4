"""The device BVH build's split order equals std::nth_element's (CPU, no GPU needed).

tests/cpp/nth_element_check.cpp compiles billardballraytracing3d_amd/csrc/bvh_select.h for the host and
compares its nth_element with the C++ library's on random, duplicate-heavy, signed-zero, sorted, reversed and
organ-pipe keys, on recursive median-split trees as BVH::build makes them (include/BVH.hpp:173-185), and on
McIlroy-adversary inputs that reach introselect's heap_select fallback.  The permutation must be the same
element for element: it is the reference tree's leaf order.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_split_order_equals_std_nth_element(tmp_path):
    exe = str(tmp_path / "nth_element_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-D__host__=", "-D__device__=", "-D__forceinline__=inline",
                    "-I", os.path.join(ROOT, "billardballraytracing3d_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "nth_element_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    fields = r.stdout.split()
    assert fields[0] == "cases" and int(fields[1]) > 3000 and fields[3] == "0", r.stdout
    assert int(fields[5]) > 0, "no input reached the heap_select fallback"
