"""BBRT_FLAG_SEED_FIRST, the network-first learned form: each candidate patch goes to Newton from the network's
seed before stage 1; stage 1 runs only when that Newton is not decisive.  Its records are measured against the
FP64 oracle, not promised equal: a patch the reference's stage 1 misses can be hit, and a farther root taken.

* the rays whose status, error kind, bounce count and first-20 patch sequence agree with the oracle's have hit
  points within 1e-10 of it, and the disagreeing rays stay a small fraction;
* the work moves as designed: stage 1 (Moller-Trumbore tests) runs for a fraction of the candidates, the
  counters add up (every candidate is decided by the network's Newton, as a hit or a miss, or by stage 1);
* the flag is refused without the learned seed.
"""
import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def disagreeing(g, r, K=20):
    kk = np.minimum(r.bounces, K)
    m = np.arange(K)[None, :] < kk[:, None]
    bad = (g.status != r.status) | (g.err_kind != r.err_kind) | (g.bounces != r.bounces) | \
        np.any(np.where(m, g.hit_patch[:, :K] != r.hit_patch[:, :K], False), 1)
    scale = np.maximum(np.linalg.norm(r.hit_xyz[:, :K], axis=-1), 1.0)
    err = np.where(m & ~bad[:, None], np.linalg.norm(g.hit_xyz[:, :K] - r.hit_xyz[:, :K], axis=-1) / scale, 0.0)
    return bad, float(err.max())


@pytest.mark.parametrize("workload", ["two_spheres", "sinai"])
def test_seed_first_against_oracle(workload):
    if workload == "two_spheres":
        sc, o, d, mb = S.two_spheres_plane_source()
    else:
        sc = S.sinai_scene()
        o, d = S.cone_rays(1 << 14, 4, 6.0, -8.0, 10.0)
        mb = 100
    g = T.GpuScene(sc)
    net = T.Mlp(11)
    rep, _ = net.train(g, n_samples=2_000_000, seed=13, batch=8192, epochs=4, lr=3e-3, lr_final=3e-5)
    g.set_mlp(net)
    gf = g.trace(o, d, mb, mode=L.NEWTON_CONVERGED, seed=L.SEED_LEARNED, flags=L.FLAG_SEED_FIRST, record_k=20)
    gl = g.trace(o, d, mb, mode=L.NEWTON_CONVERGED, seed=L.SEED_LEARNED, record_k=20)
    r = O.OracleScene(sc).trace(o, d, mb, mode=L.NEWTON_CONVERGED, record_k=20)
    bad, err = disagreeing(gf, r)
    decided = gf.counter("seed_learned") + gf.counter("seed_miss") + gf.counter("seed_fallback")
    print(f"{workload}: {o.shape[0]} rays, {int(bad.sum())} differ from the oracle ({bad.mean():.2%}), max hit error "
          f"of the others {err:.2e}; candidates {decided}: {gf.counter('seed_learned')} network hits, "
          f"{gf.counter('seed_miss')} network misses, {gf.counter('seed_fallback')} to stage 1; MT tests "
          f"{gf.counter('mt_tests')} vs {gl.counter('mt_tests')} with stage 1 first")
    assert err <= 1e-10
    assert bad.mean() <= 0.05
    assert decided > 0
    assert gf.counter("mt_tests") < gl.counter("mt_tests")
    assert gf.counter("mlp_lanes") == decided


def test_seed_first_needs_the_learned_seed():
    sc, o, d, mb = S.two_spheres_plane_source()
    g = T.GpuScene(sc)
    with pytest.raises(L.BBRTError, match="SEED_FIRST"):
        g.trace(o[:8], d[:8], mb, mode=L.NEWTON_CONVERGED, flags=L.FLAG_SEED_FIRST)
    g.set_mlp(T.Mlp(1))
    with pytest.raises(L.BBRTError, match="SEED_FIRST"):
        g.trace(o[:8], d[:8], mb, mode=L.NEWTON_CONVERGED_FP32, seed=L.SEED_LEARNED, flags=L.FLAG_SEED_FIRST)
