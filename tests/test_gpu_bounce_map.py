"""BounceMapDriver::run on the device (bbrt_bounce_map) against the oracle.

* the launch rays (grid points nudged along their normals, the theta x phi cone in the tangent frame) equal the
  oracle's restatement of BounceMapDriver.hpp:153-220 bit for bit;
* per grid point, the largest bounce count of the directions that did not end in an Error, the Error count
  and the first direction reaching the maximum equal a reduction of the oracle's traces of those rays;
* the best direction's launch ray and the grid (points, unit normals, dX/du) are consistent with the rays;
* a Sinai-scene sample at order 4 and a 180 degree cone; argument checks.
"""
import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_reduce(status, bounces, per):
    st = status.reshape(-1, per); bc = bounces.reshape(-1, per)
    ok = st != L.ERROR
    maxb = np.where(ok, bc, 0).max(1)
    errc = (~ok).sum(1)
    best = np.full(st.shape[0], -1, np.int64)
    for g in range(st.shape[0]):
        cur = -1
        for q in range(per):
            if ok[g, q] and (cur < 0 or bc[g, q] > bc[g, cur]):
                cur = q
        best[g] = cur
    return maxb, errc, best


@pytest.mark.parametrize("scene,ids,n,th,ph,cone,mb", [
    ("c1", [0, 5, 17, 300, 639], 6, 4, 5, 120.0, 20),
    ("sinai", [3, 2047, 4100, 9000], 3, 3, 3, 180.0, 30),
])
def test_bounce_map_matches_oracle(scene, ids, n, th, ph, cone, mb):
    sc = S.two_spheres_scene() if scene == "c1" else S.sinai_scene()
    gs, osc = T.GpuScene(sc), O.OracleScene(sc)
    print(f"bounce map {scene}: {len(ids)} patches, {n}x{n} points, {th}x{ph} directions", flush=True)
    g = gs.bounce_map(ids, n, th, ph, cone, mb, mode=L.NEWTON_REF, launch_rays=True)
    lo, ld = O.bounce_map_launches(osc, ids, n, th, ph, cone)
    np.testing.assert_array_equal(g["launch_o"], lo)
    np.testing.assert_array_equal(g["launch_d"], ld)
    r = osc.trace(lo, ld, mb, mode=L.NEWTON_REF)
    per = th * ph
    maxb, errc, best = oracle_reduce(r.status, r.bounces, per)
    np.testing.assert_array_equal(g["max_bounces"].ravel(), maxb)
    np.testing.assert_array_equal(g["error_count"].ravel(), errc)
    np.testing.assert_array_equal(g["best_dir"].ravel(), best)
    assert maxb.max() >= 2
    bod = g["best_od"].reshape(-1, 6)
    for q in np.flatnonzero(best >= 0):
        np.testing.assert_array_equal(bod[q, :3], lo[q * per + best[q]])
        np.testing.assert_array_equal(bod[q, 3:], ld[q * per + best[q]])
    # the grid: every launch origin is its point nudged along the (twice-normalised) unit normal by eps
    grid = g["grid"].reshape(-1, 9)
    eps = np.repeat(g["eps"], n * n)
    nrm = grid[:, 3:6] / np.linalg.norm(grid[:, 3:6], axis=1, keepdims=True)
    np.testing.assert_allclose(lo[::per], grid[:, :3] + nrm * eps[:, None], rtol=0, atol=1e-15)
    np.testing.assert_allclose(np.linalg.norm(grid[:, 3:6], axis=1), 1.0, atol=1e-15)
    # the counters are the trace's: one ray per launch
    assert g["counters"][L.CTR["rays"]] == lo.shape[0]
    assert g["counters"][L.CTR["bounces"]] == int(r.bounces.sum())


def test_bounce_map_argument_checks():
    sc = S.two_spheres_scene()
    gs = T.GpuScene(sc)
    with pytest.raises(L.BBRTError, match="grid_n"):
        gs.bounce_map([0], 1, 2, 2)
    with pytest.raises(L.BBRTError, match="out of range"):
        gs.bounce_map([640], 3, 2, 2)
    gb = T.GpuScene.batch([sc, sc])
    with pytest.raises(L.BBRTError, match="one-scene"):
        gb.bounce_map([0], 3, 2, 2)
    empty = gs.bounce_map([], 3, 2, 2)
    assert empty["max_bounces"].size == 0
