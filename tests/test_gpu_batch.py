"""Batched scenes: many sub-scenes in one handle, traced in one persistent launch.

* the batch gives, ray for ray, exactly what tracing each sub-scene alone gives (and that is the oracle's);
* hipGraph replays give the eager results every time, with eager work in between;
* the autotuner returns a launch shape whose results are the same.
"""
import numpy as np
import pytest
import torch

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("status", "err_kind", "bounces", "err_patch", "hit_patch", "hit_xyz", "iters", "final_od")


@pytest.fixture(scope="module")
def batch():
    scenes, o, d, rb = S.batch_config(seed=3, n_scenes=7, patch_lo=80, patch_hi=700, rays_lo=200, rays_hi=2500)
    return scenes, o, d, rb, T.GpuScene.batch(scenes)


def test_batch_equals_scene_by_scene(batch):
    scenes, o, d, rb, gb = batch
    mb = 30
    res = gb.trace(o, d, mb, mode=L.NEWTON_CONVERGED, ray_begin=rb, hit_count=True)
    assert res.counter("rays") == rb[-1]
    tot = np.zeros(L.NCOUNTERS, np.int64)
    for i, sc in enumerate(scenes):
        sl = slice(rb[i], rb[i + 1])
        one = T.GpuScene(sc).trace(o[sl], d[sl], mb, mode=L.NEWTON_CONVERGED, hit_count=True)
        for f in FIELDS:
            np.testing.assert_array_equal(getattr(res, f)[sl], getattr(one, f), err_msg=f"scene {i} {f}")
        np.testing.assert_array_equal(res.hit_count[gb.patch_begin[i]:gb.patch_begin[i + 1]], one.hit_count)
        tot += one.counters
    for name in ("bounces", "bvh_node_tests", "candidates", "mt_tests", "newton_calls", "newton_iters"):
        assert res.counter(name) == tot[L.CTR[name]], name
    assert res.counter("bounces") > 0


def test_batch_matches_oracle(batch):
    scenes, o, d, rb, gb = batch
    mb = 20
    res = gb.trace(o, d, mb, mode=L.NEWTON_REF, ray_begin=rb)
    for i in (0, 3):
        sl = slice(rb[i], rb[i + 1])
        r = O.OracleScene(scenes[i]).trace(o[sl], d[sl], mb, mode=L.NEWTON_REF)
        for f in ("status", "err_kind", "bounces", "iters"):
            np.testing.assert_array_equal(getattr(res, f)[sl], getattr(r, f), err_msg=f)
        m = np.arange(mb)[None, :] < r.bounces[:, None]
        np.testing.assert_array_equal(np.where(m, res.hit_patch[sl], -1), np.where(m, r.hit_patch, -1))
        np.testing.assert_array_equal(res.hit_xyz[sl][m], r.hit_xyz[m])


def _device(o, d, n, K):
    dev = torch.device("cuda", 0)
    to = torch.from_numpy(o).to(dev)
    td = torch.from_numpy(d).to(dev)
    outs = dict(status=torch.empty(n, dtype=torch.int8, device=dev),
                bounces=torch.empty(n, dtype=torch.int32, device=dev),
                hit_patch=torch.full((n, K), -7, dtype=torch.int32, device=dev),
                hit_xyz=torch.zeros((n, K, 3), dtype=torch.float64, device=dev))
    return to, td, outs


def test_graph_replays_match_eager(batch):
    scenes, o, d, rb, gb = batch
    n, K, mb = int(rb[-1]), 12, 12
    to, td, outs = _device(o, d, n, K)
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    stream = torch.cuda.current_stream().cuda_stream

    def reset():   # record slots past a ray's bounce count are never written: same sentinel every time
        outs["status"].fill_(9); outs["bounces"].fill_(-9); outs["hit_patch"].fill_(-7); outs["hit_xyz"].fill_(0.5)

    reset()
    c0, _ = gb.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, record_k=K, stream=stream, ray_begin=rb)
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in outs.items()}
    for rep in range(4):
        for v in outs.values():
            v.zero_()                      # eager work between replays
        torch.cuda.synchronize()
        reset()
        torch.cuda.synchronize()
        c, kms = gb.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, record_k=K, stream=stream,
                                 ray_begin=rb, flags=L.FLAG_GRAPH)
        torch.cuda.synchronize()
        for k in ref:
            assert torch.equal(outs[k], ref[k]), (rep, k)
        np.testing.assert_array_equal(c[:10], c0[:10])
        assert kms[2] > 0


def test_autotune_keeps_results(batch):
    scenes, o, d, rb, gb = batch
    n, K, mb = int(rb[-1]), 10, 10
    to, td, outs = _device(o, d, n, K)
    ptrs = {k: v.data_ptr() for k, v in outs.items()}
    best, rows = gb.autotune(to.data_ptr(), td.data_ptr(), n, mb, ptrs, record_k=K, ray_begin=rb)
    assert len(rows) >= 8 and best["waves_per_eu"] in (2, 3, 4) and best["newton_batch"] in (8, 16, 32, 64)
    assert min(r["ms"] for r in rows) > 0
    gb.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, record_k=K, ray_begin=rb)
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in outs.items()}
    gb.trace_device(to.data_ptr(), td.data_ptr(), n, mb, ptrs, record_k=K, ray_begin=rb, block=best["block"],
                    waves_per_eu=best["waves_per_eu"], newton_batch=best["newton_batch"],
                    grid_fraction=best["grid_fraction"])
    torch.cuda.synchronize()
    for k in ref:
        assert torch.equal(outs[k], ref[k]), k


def test_batch_handle_refuses_single_trace(batch):
    scenes, o, d, rb, gb = batch
    with pytest.raises(L.BBRTError, match="bbrt_trace_batch"):
        gb.trace(o[:5], d[:5], 5)
    bad = rb.copy()
    bad[2] = bad[1] - 1
    with pytest.raises(L.BBRTError, match="decrease"):
        gb.trace(o, d, 5, ray_begin=bad)
