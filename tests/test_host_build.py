"""The product library on the CPU: it loads, exports every symbol include/bbrt.h
declares, and its host-side scene preparation (what bbrt_scene_create uploads)
is bit-identical to the oracle's restatement of SurfacePatch's constructor and
of BVH::build; its BVH also matches the compiled reference BVH.hpp."""
import os
import ctypes as C
import re

import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "bbrt.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bbrt_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(L.EXPORTS)


def test_errors_are_reported_not_thrown():
    sc = S.two_spheres_scene()
    bad = S.subset_scene(sc, [0, 1])
    bad.imax = np.array([1, 3], np.int32)       # imax < 2
    with pytest.raises(L.BBRTError, match="imax"):
        T.prepare_patches(bad)
    with pytest.raises(L.BBRTError):
        T.build_bvh(np.zeros((0, 6)), np.zeros((0, 3)))
    # the trace entry points refuse a NULL scene before touching the GPU
    lib = L.lib()
    p, out = L.Params(), L.Outputs()
    assert lib.bbrt_trace_device_async(None, None, None, 0, None, C.byref(p), C.byref(out), None, None) == -1
    assert b"NULL scene" in lib.bbrt_last_error()
    assert lib.bbrt_trace_device(None, None, None, 0, C.byref(p), C.byref(out), None, None, None) == -1
    assert b"NULL scene" in lib.bbrt_last_error()


def _compare_with_oracle(sc, patches_to_check):
    box, cent, ms, cx, cell = T.prepare_patches(sc, coarse=True)
    osc = O.OracleScene(sc)
    obox, ocent, oms = osc.patch_info()
    np.testing.assert_array_equal(box, obox)
    np.testing.assert_array_equal(cent, ocent)
    np.testing.assert_array_equal(ms, oms)
    for k in patches_to_check:
        xyz, sag, cn = osc.coarse(k)
        np.testing.assert_array_equal(cx[k], xyz)
        np.testing.assert_array_equal(cell[k][:, 0], sag)
        np.testing.assert_array_equal(cell[k][:, 1:], cn)
    b1 = T.build_bvh(box, cent)
    b2 = osc.bvh()
    for key in b2:
        np.testing.assert_array_equal(b1[key], b2[key], err_msg=key)


def test_host_scene_matches_oracle_two_spheres():
    sc = S.two_spheres_scene()
    _compare_with_oracle(sc, [0, 1, 319, 320, 639])


def test_host_scene_matches_oracle_p4_subset():
    sc = S.sinai_scene(freq=4)     # 5 obstacles x 128 P4 patches
    _compare_with_oracle(sc, [0, 127, 128, 500, sc.n_patches - 1])


def test_host_scene_mixed_orders_and_node_sets():
    a = S.build_scene([S.Obstacle("sphere", np.zeros(3), np.ones(3), "oct", 2, 2, 1, node_kind="cheb1"),
                       S.Obstacle("ellipsoid", np.array([3.0, 0, 0]), np.array([1.5, 1, 0.6]), "ico", 1, 4, 2),
                       S.Obstacle("sphere", np.array([0, 3.0, 0]), np.ones(3), "oct", 1, 2, 3, n_nodes=6)])
    _compare_with_oracle(a, range(a.n_patches))


def test_bvh_matches_reference_harness():
    if O.ref_harness() is None:
        pytest.skip("oracle/_ref not built (no reference checkout)")
    rng = np.random.default_rng(2)
    for _ in range(50):   # random boxes and leaf thresholds: the library's tree is the oracle's
        n = int(rng.integers(1, 50))
        c = rng.uniform(-5, 5, (n, 3)); h = rng.uniform(0.05, 1.0, (n, 3))
        boxes = np.concatenate([c - h, c + h], 1)
        lt = int(rng.integers(1, 5))
        mine = T.build_bvh(boxes, c, leaf_threshold=lt)
        ref = O.bvh_build(boxes, c, leaf_threshold=lt)
        for key in ref:
            np.testing.assert_array_equal(mine[key], ref[key], err_msg=f"n={n} threshold={lt} {key}")
    # direct comparison at the scene's threshold (4)
    sc = S.two_spheres_scene()
    box, cent, _ = T.prepare_patches(sc)
    t = rng.uniform(1, 10, sc.n_patches)
    mine = T.build_bvh(box, cent, 4)
    ref = O.bvh_build(box, cent, 4)
    for key in ref:
        np.testing.assert_array_equal(mine[key], ref[key])
    for _ in range(20):
        o = rng.uniform(-4, 4, 3); d = rng.normal(size=3); d /= np.linalg.norm(d)
        a = O.bvh_visit(box, cent, t, o, d, 4)
        b = O.bvh_visit(box, cent, t, o, d, 4, fn="ref_bvh_visit")
        np.testing.assert_array_equal(a[0], b[0])


def test_noalloc_scratch_size_matches_header():
    """The Python binding's limit for bbrt_trace_device_noalloc is the header's BBRT_NOALLOC_SCRATCH_BYTES."""
    txt = open(os.path.join(ROOT, "include", "bbrt.h")).read()
    m = re.search(r"#define BBRT_NOALLOC_SCRATCH_BYTES \((\d+)u << (\d+)\)", txt)
    assert m and int(m.group(1)) << int(m.group(2)) == L.NOALLOC_SCRATCH_BYTES
    # 2^20 rays fit with the grouped queue: 60 bytes per ray + 2 * 32768 bins + alignment
    assert 60 * (1 << 20) + 8 * 32768 + 4096 <= L.NOALLOC_SCRATCH_BYTES
