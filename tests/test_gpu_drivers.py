"""The C++ drop-in drivers (include/bbrt/drivers.hpp) end to end on the GPU.

tests/cpp/drivers_check runs RayTracingDriver / PlaneSourceDriver /
BounceMapDriver on a patch directory written in the reference's file format;
the results are compared with the CPU oracle run on the same patches in the
driver's patch order (the lexicographic file order of
RayTracingDriver.hpp:361), so that the BVH, and with it the candidate visit
order and every error event, is the same.
"""
import os
import subprocess

import numpy as np
import pytest

from billardballraytracing3d_amd import scenes as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "drivers_check")


def run(args):
    print("[drivers]", " ".join(str(a) for a in args[:2]), flush=True)
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr


def load(path, dtype):
    return np.fromfile(path, dtype=dtype)


@pytest.fixture(scope="module")
def c1dir(tmp_path_factory):
    sc, o, d, mb = S.two_spheres_plane_source()
    d_ = tmp_path_factory.mktemp("two_spheres")
    S.write_patch_dir(sc, str(d_))
    return sc, o, d, mb, str(d_) + "/"


def driver_order_scene(sc, zone):
    # zoneID = python patch index + 1 (write_patch_dir); the driver sorted the files
    return S.subset_scene(sc, (zone - 1).tolist())


def test_ray_tracing_driver_matches_oracle(c1dir, tmp_path):
    from oracle import oracle as O
    sc, o, d, mb, dirp = c1dir
    rays = np.ascontiguousarray(np.concatenate([o, d], 1))
    rays[::97, 3:] = 0.0   # a few zero directions: set_rays rejects them, so drop those rays
    keep = np.any(rays[:, 3:] != 0, axis=1)
    rays = np.ascontiguousarray(rays[keep])
    rays.tofile(tmp_path / "rays.bin")
    out = str(tmp_path / "t")
    run(["trace", dirp, tmp_path / "rays.bin", rays.shape[0], mb, 0, out])
    zone = load(out + ".zone", np.int64)
    assert sorted(zone.tolist()) == list(range(1, sc.n_patches + 1))
    osc = O.OracleScene(driver_order_scene(sc, zone))
    r = osc.trace(rays[:, :3], rays[:, 3:], mb, mode=0)
    n = rays.shape[0]
    status = load(out + ".status", np.int8)
    bounces = load(out + ".bounces", np.int32)
    hits = load(out + ".hits", np.int32).reshape(n, mb)
    pts = load(out + ".pts", np.float64).reshape(n, mb + 2, 3)
    np.testing.assert_array_equal(status, r.status)
    np.testing.assert_array_equal(bounces, r.bounces)
    np.testing.assert_array_equal(hits, r.hit_patch)
    for i in range(n):
        b = r.bounces[i]
        np.testing.assert_array_equal(pts[i, 0], rays[i, :3])
        np.testing.assert_array_equal(pts[i, 1:1 + b], r.hit_xyz[i, :b])
        if r.status[i] == 0:   # escaped: tail point = last origin + dir * scene diagonal
            tail = r.final_od[i, :3] + r.final_od[i, 3:] * osc.diag
            np.testing.assert_array_equal(pts[i, 1 + b], tail)
    errs = open(out + ".err").read().split("\n")
    for i in np.where(r.status == 2)[0]:
        assert errs[i].startswith({1: "SurfacePatch::intersect: singular", 2: "SurfacePatch::intersect: Newton refinement failed",
                                   3: "SurfacePatch::intersect: Newton refinement converged outside",
                                   4: "RayTracingDriver::trace_ray: hit back face"}[int(r.err_kind[i])])
    assert os.path.getsize(out + ".vtk") > 0


@pytest.mark.parametrize("count_direct", [0, 1])
def test_plane_source_driver_hit_counts(c1dir, tmp_path, count_direct):
    from oracle import oracle as O
    sc, o, d, mb, dirp = c1dir
    out = str(tmp_path / "p")
    run(["plane", dirp, -3, -3, -5, 3, 3, -5, 0, 0, 1, 64, 64, mb, count_direct, out])
    zone = load(out + ".zone", np.int64)
    hc = load(out + ".hitcount", np.int64)
    osc = O.OracleScene(driver_order_scene(sc, zone))
    grid = S.plane_grid((-3, -3, -5), (3, 3, -5), 64, 64)
    r = osc.trace(grid, np.tile([0.0, 0.0, 1.0], (grid.shape[0], 1)), mb, mode=0)
    want = np.zeros(sc.n_patches, np.int64)
    for i in range(grid.shape[0]):
        for k in range(0 if count_direct else 1, r.bounces[i]):
            want[r.hit_patch[i, k]] += 1
    np.testing.assert_array_equal(hc, want)
    assert want.sum() > 0


def test_bounce_map_driver(c1dir, tmp_path):
    from oracle import oracle as O
    sc, o, d, mb, dirp = c1dir
    out = str(tmp_path / "b")
    first, last, n, th, ph, cone = 5, 12, 4, 3, 4, 120.0
    run(["bmap", dirp, "patch", first, last, n, th, ph, cone, mb, 1, out])
    zone = load(out + ".zone", np.int64)
    osc = O.OracleScene(driver_order_scene(sc, zone))
    pos = {int(z): k for k, z in enumerate(zone)}
    ids = [pos[i] for i in range(first, last + 1)]      # file prefix<i>.txt has zoneID i
    lo, ld = O.bounce_map_launches(osc, ids, n, th, ph, cone)
    np.testing.assert_array_equal(load(out + ".launch_o", np.float64).reshape(-1, 3), lo)
    np.testing.assert_array_equal(load(out + ".launch_d", np.float64).reshape(-1, 3), ld)
    r = osc.trace(lo, ld, mb, mode=0)
    per = th * ph
    npts = (last - first + 1) * n * n
    st = r.status.reshape(npts, per)
    bc = r.bounces.reshape(npts, per)
    want_max = np.where(st != 2, bc, 0).max(1)
    want_err = (st == 2).sum(1)
    np.testing.assert_array_equal(load(out + ".maxb", np.int32), want_max)
    np.testing.assert_array_equal(load(out + ".errc", np.int32), want_err)
    best = load(out + ".bestb", np.int32)
    has = (st != 2).any(1)
    np.testing.assert_array_equal(best[has], want_max[has])
    assert os.path.isdir(out + "_best")
