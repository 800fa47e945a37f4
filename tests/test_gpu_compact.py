"""Compact scenes (BBRT_SCENE_COMPACT): coarse cells evaluated on the fly from the nodal fields.

The stored and compact scenes must give identical records and identical work counters: the on-the-fly
corners, sagittas, normals and FP32 cell boxes are bit-identical to the stored ones, so the same cells pass
the same tests.
"""
import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T

pytestmark = pytest.mark.gpu

FIELDS = ("status", "err_kind", "bounces", "err_patch", "hit_patch", "hit_xyz", "iters", "final_od")


def _same(a, b):
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    np.testing.assert_array_equal(a.counters[:10], b.counters[:10])


@pytest.mark.parametrize("mode", [L.NEWTON_REF, L.NEWTON_CONVERGED])
def test_compact_equals_stored_two_spheres(mode):
    sc, o, d, mb = S.two_spheres_plane_source()
    st = T.GpuScene(sc)
    cp = T.GpuScene.batch([sc], compact=True)
    _same(cp.trace(o, d, mb, mode=mode), st.trace(o, d, mb, mode=mode))
    assert cp.info()["device_bytes"] * 5 < st.info()["device_bytes"]


def test_compact_equals_stored_sinai_sample():
    sc, o, d, mb = S.sinai_cone(0, 1 << 20)
    sel = slice(0, 1 << 20, 512)
    st = T.GpuScene(sc)
    cp = T.GpuScene.batch([sc], compact=True)
    a = cp.trace(o[sel], d[sel], mb, mode=L.NEWTON_CONVERGED)
    _same(a, st.trace(o[sel], d[sel], mb, mode=L.NEWTON_CONVERGED))
    assert a.counter("bounces") > 0


def test_compact_batch_equals_stored_batch():
    scenes, o, d, rb = S.batch_config(seed=5, n_scenes=5, patch_lo=80, patch_hi=600, rays_lo=300, rays_hi=2000)
    a = T.GpuScene.batch(scenes, compact=True).trace(o, d, 25, ray_begin=rb)
    b = T.GpuScene.batch(scenes).trace(o, d, 25, ray_begin=rb)
    _same(a, b)


def test_compact_brute_force_equals_stored():
    sc, o, d, mb = S.two_spheres_plane_source(nx=8, ny=8)
    st = T.GpuScene(sc)
    cp = T.GpuScene.batch([sc], compact=True)
    _same(cp.trace(o, d, 3, flags=L.FLAG_NO_CULL), st.trace(o, d, 3, flags=L.FLAG_NO_CULL))


def test_compact_scene_has_no_coarse_grid():
    sc, _, _, _ = S.two_spheres_plane_source(nx=2, ny=2)
    cp = T.GpuScene.batch([sc], compact=True)
    with pytest.raises(L.BBRTError, match="compact"):
        cp.coarse(0)


def test_device_build_equals_host_build():
    """The device build of a compact scene gives the host build's boxes, BVH and, so, records and counters."""
    scenes, o, d, rb = S.batch_config(seed=6, n_scenes=4, patch_lo=80, patch_hi=900, rays_lo=300, rays_hi=3000)
    a = T.GpuScene.batch(scenes, compact=True, device_build=True)
    b = T.GpuScene.batch(scenes, compact=True)
    for x, y in zip(a.patch_info(), b.patch_info()):
        np.testing.assert_array_equal(x, y)
    ba, bb = a.bvh(), b.bvh()
    for key in ba:
        np.testing.assert_array_equal(ba[key], bb[key], err_msg=key)
    _same(a.trace(o, d, 25, ray_begin=rb), b.trace(o, d, 25, ray_begin=rb))
    sc, o1, d1, mb = S.two_spheres_plane_source()
    _same(T.GpuScene.batch([sc], compact=True, device_build=True).trace(o1, d1, mb), T.GpuScene(sc).trace(o1, d1, mb))



def test_wave_cell_tests_on_mixed_orders_and_brute_force():
    """Compact stage 1 tests the cells the lanes' walks select wave-wide: a lane's cells can be tested by
    other lanes, for patches of other orders, and a brute-force walk (every tile) spans many passes of the
    wave's list.  Records and work counters equal the stored scene's, culled and brute force, with the
    triangle and with the learned seed (whose count of wave-level network passes follows the schedule)."""
    scenes, o, d, rb = S.batch_config(seed=8, n_scenes=4, patch_lo=80, patch_hi=700, rays_lo=400, rays_hi=2500)
    cp = T.GpuScene.batch(scenes, compact=True, device_build=True)
    st = T.GpuScene.batch(scenes)
    keep = np.arange(L.NCOUNTERS) != L.CTR["mlp_evals"]
    for learned in (False, True):
        seed = L.SEED_TRIANGLE
        if learned:
            net = T.Mlp(4)
            cp.set_mlp(net)
            st.set_mlp(net)
            seed = L.SEED_LEARNED
        for fl in (0, L.FLAG_NO_CULL):
            a = cp.trace(o, d, 25, ray_begin=rb, mode=L.NEWTON_CONVERGED, seed=seed, flags=fl)
            b = st.trace(o, d, 25, ray_begin=rb, mode=L.NEWTON_CONVERGED, seed=seed, flags=fl)
            _same(a, b)
            np.testing.assert_array_equal(a.counters[keep], b.counters[keep])
            assert a.counter("bounces") > 0
