"""The BVH built on the GPU is the reference's tree, bit for bit.

bbrt_build_bvh_gpu runs the device build (billardballraytracing3d_amd/csrc/bbrt_bvh.hip) over arbitrary objects;
bbrt_build_bvh runs BVH::build's steps on the host with std::nth_element (include/BVH.hpp:136-189).  Every node
box, child link, leaf range, split axis and the leaf order must be equal, on inputs chosen to break a plausible
mistake: equal centroids (the split order is then the library's permutation, not a sort), few distinct values,
signed zeros and NaN in the boxes (the folds' first-occurrence rules), node ranges longer than the LDS stage (the
device-memory phase of the select), leaf thresholds 1 to 7, and a McIlroy-style input whose splits reach
introselect's heap_select fallback.  Scene handles built on the device (BBRT_SCENE_DEVICE_BUILD) get the host
build's trees for a 131-scene batch, and trace to the same records.
"""
import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T

pytestmark = pytest.mark.gpu


def _boxes_around(cent, rng, half=0.05):
    h = rng.uniform(0.2, 1.0, cent.shape) * half
    return np.concatenate([cent - h, cent + h], axis=1)


def _same_tree(box, cent, leaf_threshold):
    a = T.build_bvh(box, cent, leaf_threshold)
    b = T.build_bvh(box, cent, leaf_threshold, gpu=True)
    for k in ("right", "start", "count", "axis", "leaf"):
        np.testing.assert_array_equal(b[k], a[k], err_msg=k)
    # boxes bit for bit (NaN payloads and the sign of zero included)
    np.testing.assert_array_equal(b["box"].view(np.uint64), a["box"].view(np.uint64))
    return a


@pytest.mark.parametrize("n,leaf_threshold", [(1, 4), (5, 4), (37, 1), (1000, 2), (4099, 4), (9000, 7)])
def test_random_objects(n, leaf_threshold):
    rng = np.random.default_rng(n)
    cent = rng.uniform(-3, 3, (n, 3))
    t = _same_tree(_boxes_around(cent, rng), cent, leaf_threshold)
    assert len(t["right"]) >= 1


def test_equal_and_few_distinct_centroids():
    rng = np.random.default_rng(7)
    n = 7000   # the top splits run past the LDS stage
    cent = np.zeros((n, 3))
    _same_tree(_boxes_around(cent, rng), cent, 4)                     # all equal: the library's order decides
    cent = rng.integers(0, 3, (n, 3)).astype(np.float64) * 0.5         # 27 distinct points
    _same_tree(_boxes_around(cent, rng), cent, 4)
    cent = np.round(rng.uniform(-1, 1, (n, 3)), 1)                    # a lattice: ties on every axis
    cent[:, 2] = 0.0
    _same_tree(_boxes_around(cent, rng), cent, 3)


def test_signed_zeros_and_nan_boxes():
    rng = np.random.default_rng(11)
    n = 600
    cent = rng.uniform(-1, 1, (n, 3))
    box = _boxes_around(cent, rng)
    box[rng.integers(0, n, 80), rng.integers(0, 6, 80)] = 0.0
    box[rng.integers(0, n, 80), rng.integers(0, 6, 80)] = -0.0
    box[rng.integers(0, n, 5), rng.integers(0, 6, 5)] = np.nan
    box[0, 1] = np.nan    # a NaN at the very first object: the fold keeps it
    cent[rng.integers(0, n, 40), 0] = -0.0
    _same_tree(box, cent, 4)


def test_heap_select_fallback_input():
    """Keys from McIlroy's adversary against the median-of-three partition (the same construction as
    tests/cpp/nth_element_check.cpp): the root split runs out of partition rounds and heap-selects."""
    n = 2000
    val = np.full(n, n, np.int64)
    state = dict(solid=0, cand=0)

    def less(x, y):
        if val[x] == n and val[y] == n:
            z = x if x == state["cand"] else y
            val[z] = state["solid"]; state["solid"] += 1
        if val[x] == n:
            state["cand"] = x
        elif val[y] == n:
            state["cand"] = y
        return val[x] < val[y]

    # introselect's steps over ids with the adversary's comparisons (python, one pass)
    a = list(range(n))
    lo, hi, nth = 0, n, n // 2
    depth = 2 * (n.bit_length() - 1)
    rounds = 0
    while hi - lo > 3 and depth > 0:
        depth -= 1
        rounds += 1
        mid = lo + (hi - lo) // 2
        x, y, z = lo + 1, mid, hi - 1
        if less(a[x], a[y]):
            m = y if less(a[y], a[z]) else (z if less(a[x], a[z]) else x)
        else:
            m = x if less(a[x], a[z]) else (z if less(a[y], a[z]) else y)
        a[lo], a[m] = a[m], a[lo]
        f, l, pv = lo + 1, hi, lo
        while True:
            while less(a[f], a[pv]):
                f += 1
            l -= 1
            while less(a[pv], a[l]):
                l -= 1
            if not f < l:
                break
            a[f], a[l] = a[l], a[f]
            f += 1
        if f <= nth:
            lo = f
        else:
            hi = f
    assert depth == 0, "the adversary did not exhaust the partition rounds"
    cent = np.zeros((n, 3))
    cent[:, 0] = val.astype(np.float64)
    cent[:, 1] = np.linspace(0, 1e-3, n)   # x stays the widest axis
    _same_tree(_boxes_around(cent, np.random.default_rng(3)), cent, 4)


def test_device_built_batch_trees_equal_host_built():
    scenes, o, d, rb = S.batch_config(seed=5, n_scenes=131, patch_lo=80, patch_hi=100, rays_lo=20, rays_hi=60)
    dev = T.GpuScene.batch(scenes, compact=True, device_build=True)
    host = T.GpuScene.batch(scenes, compact=True)
    ta, tb = dev.bvh(), host.bvh()
    for k in tb:
        np.testing.assert_array_equal(ta[k], tb[k], err_msg=k)
    assert dev.info()["n_nodes"] == host.info()["n_nodes"]
    a = dev.trace(o, d, 25, ray_begin=rb)
    b = host.trace(o, d, 25, ray_begin=rb)
    for f in ("status", "err_kind", "bounces", "err_patch", "hit_patch", "hit_xyz", "iters", "final_od"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    np.testing.assert_array_equal(a.counters[:10], b.counters[:10])
