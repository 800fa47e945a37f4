"""A batch handle with many small sub-scenes (131 scenes, compact, built on the device).

* The grouped ray queue has 131 x 4^3 = 8,384 bins: more than one 8,192-bin pass of k_scan_bins and not a
  multiple of it.  Records and counters equal the input-order queue's, and sampled sub-scenes equal the oracle.
"""
import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("status", "err_kind", "bounces", "err_patch", "hit_patch", "hit_xyz", "iters", "final_od")


@pytest.fixture(scope="module")
def many():
    scenes, o, d, rb = S.batch_config(seed=5, n_scenes=131, patch_lo=80, patch_hi=100, rays_lo=20, rays_hi=60)
    return scenes, o, d, rb, T.GpuScene.batch(scenes, compact=True, device_build=True)


def test_grouped_queue_over_several_scan_passes(many):
    scenes, o, d, rb, gb = many
    mb = 25
    base = gb.trace(o, d, mb, ray_begin=rb, hit_count=True)
    grp = gb.trace(o, d, mb, ray_begin=rb, hit_count=True, flags=L.FLAG_GROUP_RAYS)
    for f in FIELDS:
        np.testing.assert_array_equal(getattr(grp, f), getattr(base, f), err_msg=f)
    np.testing.assert_array_equal(grp.hit_count, base.hit_count)
    np.testing.assert_array_equal(grp.counters[:10], base.counters[:10])
    assert base.counter("bounces") > 0
    for i in (0, 64, 130):
        sl = slice(rb[i], rb[i + 1])
        r = O.OracleScene(scenes[i]).trace(o[sl], d[sl], mb, mode=L.NEWTON_CONVERGED)
        for f in ("status", "err_kind", "bounces", "iters"):
            np.testing.assert_array_equal(getattr(grp, f)[sl], getattr(r, f), err_msg=f"scene {i} {f}")
        m = np.arange(mb)[None, :] < r.bounces[:, None]
        np.testing.assert_array_equal(np.where(m, grp.hit_patch[sl], -1), np.where(m, r.hit_patch, -1))
        assert np.abs(grp.hit_xyz[sl][m] - r.hit_xyz[m]).max(initial=0.0) <= 1e-10
