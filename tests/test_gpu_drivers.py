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


def read_vtk(path):
    """Legacy ASCII VTK POLYDATA as the drivers write it (RayTracingDriver.hpp:270-311,
    PlaneSourceDriver.hpp:115-185, BounceMapDriver.hpp:267-408): header, POINTS, LINES or POLYGONS,
    CELL_DATA / POINT_DATA scalars."""
    tok = open(path).read().split("\n")
    assert tok[0] == "# vtk DataFile Version 3.0" and tok[2] == "ASCII" and tok[3] == "DATASET POLYDATA"
    out = {"title": tok[1], "scalars": {}}
    i = 4
    while i < len(tok):
        line = tok[i].split()
        i += 1
        if not line:
            continue
        if line[0] == "POINTS":
            n = int(line[1])
            assert line[2] == "double"
            out["points"] = np.array([[float(v) for v in tok[i + k].split()] for k in range(n)]).reshape(n, 3)
            i += n
        elif line[0] in ("LINES", "POLYGONS"):
            n, size = int(line[1]), int(line[2])
            cells = [[int(v) for v in tok[i + k].split()] for k in range(n)]
            assert sum(len(c) for c in cells) == size
            assert all(c[0] == len(c) - 1 for c in cells)
            out[line[0].lower()] = np.array([c[1:] for c in cells], np.int64).reshape(n, -1)
            i += n
        elif line[0] in ("CELL_DATA", "POINT_DATA"):
            out[line[0].lower()] = int(line[1])
        elif line[0] == "SCALARS":
            assert tok[i] == "LOOKUP_TABLE default"
            n = out.get("point_data") if "point_data" in out else out["cell_data"]
            out["scalars"][line[1]] = np.array([int(tok[i + 1 + k]) for k in range(n)], np.int64)
            i += 1 + n
    return out


def check_path_vtk(v, paths_pts, id_name):
    """paths_pts: per path, its points (origin, hits, escape tail).  One 2-point LINE per leg, with the path
    index and the leg index as scalars; points bit for bit."""
    want = np.concatenate(paths_pts) if paths_pts else np.zeros((0, 3))
    assert v["points"].shape == want.shape
    assert np.array_equal(v["points"].view(np.uint64), want.view(np.uint64))
    segs, ids, ks = [], [], []
    base = 0
    for i, pp in enumerate(paths_pts):
        for k in range(len(pp) - 1):
            segs.append((base + k, base + k + 1)); ids.append(i); ks.append(k)
        base += len(pp)
    np.testing.assert_array_equal(v["lines"], np.array(segs, np.int64).reshape(-1, 2))
    np.testing.assert_array_equal(v["scalars"][id_name], ids)
    np.testing.assert_array_equal(v["scalars"]["BounceIndex"], ks)


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
    # the VTK file holds every path's points in ray order, one line per leg
    paths_pts = []
    for i in range(n):
        m = 1 + r.bounces[i] + (1 if r.status[i] == 0 else 0)
        paths_pts.append(pts[i, :m])
    v = read_vtk(out + ".vtk")
    assert v["title"] == "Billiard-ball ray traces"
    check_path_vtk(v, paths_pts, "RayID")


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
    # the VTK file: every patch's nodes in driver order, two triangles per node cell, the patch's count per node
    v = read_vtk(out + ".vtk")
    assert v["title"] == "Plane-source patch hit counts"
    sco = driver_order_scene(sc, zone)
    nodes, tris, cnt = [], [], []
    base = 0
    for k in range(sco.n_patches):
        F = sco.patch_fields(k)
        im, jm = int(sco.imax[k]), int(sco.jmax[k])
        nodes.append(F[:3].T)
        cnt += [want[k]] * (im * jm)
        for i in range(im - 1):
            for j in range(jm - 1):
                p00, p10 = base + i * jm + j, base + (i + 1) * jm + j
                p01, p11 = base + i * jm + j + 1, base + (i + 1) * jm + j + 1
                tris += [(p00, p10, p11), (p00, p11, p01)]
        base += im * jm
    want_pts = np.concatenate(nodes)
    assert np.array_equal(v["points"].view(np.uint64), want_pts.view(np.uint64))
    np.testing.assert_array_equal(v["polygons"], np.array(tris))
    np.testing.assert_array_equal(v["scalars"]["HitCount"], cnt)


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
    # the bounce-map VTK: each sampled patch's n x n grid (the launch points), two triangles per grid cell, the
    # per-point maximum bounce count and error count
    v = read_vtk(out + ".vtk")
    assert v["title"] == "Billiard-ball bounce map"
    assert v["points"].shape == (npts, 3)
    scale = osc.patch_info()[2][ids].max()
    assert np.abs(v["points"] - lo[::per]).max() <= 10 * scale + 1e-9 * osc.diag   # launch = point + nudge
    np.testing.assert_array_equal(v["scalars"]["MaxBounces"], want_max)
    np.testing.assert_array_equal(v["scalars"]["ErrorCount"], want_err)
    tris = []
    for p in range(last - first + 1):
        for i in range(n - 1):
            for j in range(n - 1):
                b0 = p * n * n
                tris += [(b0 + i * n + j, b0 + (i + 1) * n + j, b0 + (i + 1) * n + j + 1),
                         (b0 + i * n + j, b0 + (i + 1) * n + j + 1, b0 + i * n + j + 1)]
    np.testing.assert_array_equal(v["polygons"], np.array(tris))
    # the best-ray files: one per sampled patch, one path per grid point that has a non-error ray, starting at
    # that point's launch origin, with the point's maximum bounce count
    files = sorted(os.listdir(out + "_best"))
    assert files == sorted(f"best_rays_patch_{i}.vtk" for i in range(first, last + 1))
    for p in range(last - first + 1):
        bv = read_vtk(os.path.join(out + "_best", f"best_rays_patch_{first + p}.vtk"))
        pid = bv["scalars"]["PointID"]
        assert bv["lines"].shape[0] == pid.size
        starts = [0] + [int(b) + 1 for b in np.flatnonzero(np.diff(pid) != 0)]
        npath = len(np.unique(pid))
        assert npath <= n * n and npath == len(starts)
        for q in range(npath):
            seg = bv["lines"][starts[q]]
            first_pt = bv["points"][seg[0]]
            g = p * n * n + int(pid[starts[q]])
            np.testing.assert_array_equal(first_pt, lo[g * per])
            nseg = int((pid == pid[starts[q]]).sum())
            if has[g]:
                assert nseg in (best[g], best[g] + 1)   # hits, plus the escape leg
