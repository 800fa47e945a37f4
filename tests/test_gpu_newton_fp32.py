"""BBRT_NEWTON_CONVERGED_FP32 (Newton's first iterations in FP32, the reference's stop, determinant and domain
decisions and the polish in FP64) against the FP64 oracle in converged mode.

Its classification is measured, not promised: a ray whose FP32 and FP64 iterations part ways can end
differently.  The tests bound that rate on the two-sphere plane source, a Sinai sample and compact batched
scenes of the full form's layout, check the hit points of every ray whose records agree to 1e-10, and check
that the FP32 phase really carries the iterations (FP64 iterations fall, FP32 iterations are counted).
"""
from types import SimpleNamespace

import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def compare(g, r, K=20):
    """(rays whose status, error kind, bounce count or first-K patch sequence differ, max relative hit error
    over the others)."""
    kk = np.minimum(r.bounces, K)
    m = np.arange(g.hit_patch.shape[1])[None, :K] < kk[:, None]
    bad = (g.status != r.status) | (g.err_kind != r.err_kind) | (g.bounces != r.bounces) | \
        np.any(np.where(m, g.hit_patch[:, :K] != r.hit_patch[:, :K], False), 1)
    scale = np.maximum(np.linalg.norm(r.hit_xyz[:, :K], axis=-1), 1.0)
    err = np.where(m & ~bad[:, None], np.linalg.norm(g.hit_xyz[:, :K] - r.hit_xyz[:, :K], axis=-1) / scale, 0.0)
    return int(bad.sum()), float(err.max())


@pytest.mark.parametrize("workload", ["two_spheres", "sinai"])
def test_fp32_start_against_fp64_oracle(workload):
    if workload == "two_spheres":
        sc, o, d, mb = S.two_spheres_plane_source()
    else:
        sc, o, d, mb = S.sinai_cone(n_rays=1 << 13)
        o, d, mb = o[:3000], d[:3000], 100
    gs, osc = T.GpuScene(sc), O.OracleScene(sc)
    g32 = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED_FP32)
    g64 = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    r = osc.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    bad, err = compare(g32, r)
    print(f"{workload}: {o.shape[0]} rays, {bad} with a changed classification or sequence, max hit error {err:.2e}; "
          f"FP64 iterations {g32.counter('newton_iters')} vs {g64.counter('newton_iters')}, FP32 iterations "
          f"{g32.counter('newton_iters_fp32')}")
    assert bad <= max(2, o.shape[0] // 200)
    assert err <= 1e-10
    assert g32.counter("newton_iters_fp32") > 0
    assert g32.counter("newton_iters") < g64.counter("newton_iters")
    assert g64.counter("newton_iters_fp32") == 0


def test_fp32_start_on_compact_batched_scenes():
    scenes, o, d, rb = S.batch_config(seed=9, n_scenes=4, patch_lo=80, patch_hi=800, rays_lo=400, rays_hi=1500)
    gb = T.GpuScene.batch(scenes, compact=True, device_build=True)
    mb = 50
    g = gb.trace(o, d, mb, mode=L.NEWTON_CONVERGED_FP32, ray_begin=rb)
    total_bad = total = 0
    for i in range(len(scenes)):
        sl = slice(int(rb[i]), int(rb[i + 1]))
        r = O.OracleScene(scenes[i]).trace(o[sl], d[sl], mb, mode=L.NEWTON_CONVERGED)
        part = SimpleNamespace(**{f: getattr(g, f)[sl] for f in ("status", "err_kind", "bounces", "hit_patch", "hit_xyz")})
        bad, err = compare(part, r)
        total_bad += bad
        total += r.status.size
        assert err <= 1e-10
    print(f"compact batch: {total} rays, {total_bad} with a changed classification or sequence")
    assert total_bad <= max(2, total // 200)
    assert g.counter("newton_iters_fp32") > 0


def test_fp32_start_with_learned_seeds():
    sc, o, d, mb = S.two_spheres_plane_source()
    gs = T.GpuScene(sc)
    gs.set_mlp(T.Mlp(3))
    gl = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED_FP32, seed=L.SEED_LEARNED)
    g32 = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED_FP32)
    differ = int(((gl.status != g32.status) | (gl.bounces != g32.bounces)).sum())
    assert differ <= max(2, o.shape[0] // 200)
    assert gl.counter("seed_learned") + gl.counter("seed_fallback") > 0
    with pytest.raises(L.BBRTError, match="newton_mode"):
        gs.trace(o[:4], d[:4], mb, mode=3)
