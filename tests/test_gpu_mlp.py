"""The learned Newton seed on the GPU, against the oracle (oracle/mlp_oracle.py, oracle/bbrt_oracle.cpp).

* forward pass and gradients of the bf16 matrix-core kernels vs the oracle's emulation of the same roundings;
* features and training targets vs the oracle, bit for bit;
* training lowers the loss, and the trained seed's Newton lands on the triangle seed's root;
* the bounce loop with the learned seed gives the triangle-seed results (patch sequences exact, points within
  1e-10) -- with a trained network and with an untrained one, whose seeds all fall back.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _params(seed, scale):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(M.N_PARAMS) * scale).astype(np.float32) * M.live_mask()


def test_forward_matches_oracle():
    net = T.Mlp(0)
    p = _params(11, 0.25)
    net.params = p
    rng = np.random.default_rng(12)
    x = rng.standard_normal((1000, M.N_IN)).astype(np.float32)   # 15 full tiles + a ragged one
    y = net.forward(x)
    want, _ = M.forward(p, x, bf16_ops=True)
    err = np.abs(y - want)
    scale = np.abs(want).max()
    # fp32 accumulation order differs; a rare bf16 rounding flip of a hidden value moves an output slightly
    assert np.median(err) <= 1e-5 * scale, np.median(err)
    assert np.mean(err > 1e-3 * scale) < 0.01
    assert err.max() <= 3e-2 * scale, err.max()
    # and far from the unrounded network (the roundings are really applied)
    exact, _ = M.forward(p, x, bf16_ops=False)
    assert np.abs(exact - want).max() > 10 * np.median(err)


def test_gradient_matches_oracle():
    net = T.Mlp(0)
    p = _params(13, 0.2)
    net.params = p
    rng = np.random.default_rng(14)
    n = 293                                                        # ragged last tile
    x = rng.standard_normal((n, M.N_IN)).astype(np.float32)
    t = rng.standard_normal((n, 3)).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    lo, go = M.loss_grad(p, x, t, bf16_ops=True)
    assert abs(loss - lo) <= 1e-3 * lo
    live = M.live_mask() > 0
    assert np.all(g[~live] == 0)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    assert rel < 1e-2, rel
    # every block separately (a transposed or dropped block would show here)
    offs = [M.OFF_W1, M.OFF_B1, M.OFF_W2, M.OFF_B2, M.OFF_W3, M.OFF_B3, M.OFF_W4, M.OFF_B4, M.N_PARAMS]
    for a, b in zip(offs[:-1], offs[1:]):
        ga, gb = g[a:b], go[a:b]
        assert np.linalg.norm(ga - gb) <= 2e-2 * np.linalg.norm(gb) + 1e-7, (a, np.linalg.norm(ga - gb), np.linalg.norm(gb))
    # one Adam-free step along -g lowers the loss (the gradient is a descent direction)
    net.params = p - 1e-3 * g * M.live_mask()
    l2, _ = net.loss_grad(x, t)
    assert l2 < loss


@pytest.fixture(scope="module")
def c1():
    sc, o, d, mb = S.two_spheres_plane_source()
    return sc, T.GpuScene(sc), O.OracleScene(sc), o, d, mb


def test_features_and_targets_match_oracle(c1):
    sc, g, osc, _, _, _ = c1
    patch, o, d, x, t, valid = g.sample_pairs(seed=5, first=1000, n=300)
    assert valid.mean() > 0.9
    x2, tc, frame = g.mlp_features(patch, o, d, frame=True)
    np.testing.assert_array_equal(x2, x)
    for i in range(0, 300, 7):
        fr = M.scene_patch_frame(sc, int(patch[i]))
        np.testing.assert_array_equal(frame[patch[i], :3], fr[0])
        np.testing.assert_array_equal(frame[patch[i], 3:6], fr[1])
        np.testing.assert_array_equal(frame[patch[i], 12], fr[4])
        np.testing.assert_array_equal(frame[patch[i], 16:19], fr[6])
        np.testing.assert_array_equal(frame[patch[i], 19:22], fr[7])
        feats = M.ray_features(fr, o[i], d[i])
        np.testing.assert_array_equal(x[i, :6], feats[0])
        np.testing.assert_array_equal(x[i, 6:], fr[5])
        np.testing.assert_array_equal(tc[i], feats[1:])
        st, tt, _, _, _, u, v = osc.patch_intersect_uv(int(patch[i]), o[i], d[i], mode=1)
        assert valid[i] == (st == 1)
        if st == 1:
            np.testing.assert_array_equal(t[i], M.target_from_root(fr, feats, u, v, tt))


@pytest.fixture(scope="module")
def trained(c1):
    sc, g, _, _, _, _ = c1
    net = T.Mlp(3)
    rep, curve = net.train(g, n_samples=2_000_000, seed=21, batch=4096, epochs=8, lr=2e-3, lr_final=2e-5)
    g.set_mlp(net)
    return net, rep, curve


def test_training_lowers_loss(trained):
    net, rep, curve = trained
    assert rep["n_pairs"] > 0.9 * 2_000_000
    assert rep["steps"] == len(curve) > 1000
    assert curve[-50:].mean() < 0.25 * curve[:5].mean()
    assert rep["val_loss"] < 2.0 * rep["final_loss"] + 1e-4
    assert np.isfinite(net.params).all()


def test_training_leaves_fragments_equal_to_a_fresh_pack(trained):
    """Each Adam step writes the updated weights straight into the bf16 fragments (forward, transposed and
    ray fragments).  After training they must be what packing the downloaded masters afresh gives: the
    forward pass and the gradient (which reads the transposed fragments) agree bit for bit."""
    net, _, _ = trained
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((3000, M.N_IN)) * 0.5).astype(np.float32)
    t = (rng.standard_normal((3000, 3)) * 0.1).astype(np.float32)
    y1 = net.forward(x)
    l1, g1 = net.loss_grad(x, t)
    twin = T.Mlp(0)
    twin.params = net.params
    y2 = twin.forward(x)
    l2, g2 = twin.loss_grad(x, t)
    np.testing.assert_array_equal(y1, y2)
    np.testing.assert_array_equal(g1, g2)
    assert l1 == l2


def test_learned_seed_lands_on_the_triangle_root(c1, trained):
    _, g, _, _, _, _ = c1
    ab = g.seed_ab(seed=99, first=0, n=20000)           # pairs never used in training (other seed)
    hit = ab["status_tri"] == 1
    acc = ab["accepted"] & hit
    assert acc.mean() / hit.mean() > 0.9
    assert np.all(ab["point_diff"][acc] <= 1e-10)
    # the network's seeds are better than chance: well inside a patch
    assert np.median(ab["seed_err_learned"][hit]) < 0.05


def _compare(gr, orr, bounces=20):
    np.testing.assert_array_equal(gr.status, orr.status)
    np.testing.assert_array_equal(gr.err_kind, orr.err_kind)
    np.testing.assert_array_equal(gr.bounces, orr.bounces)
    k = min(bounces, gr.hit_patch.shape[1])
    for i in np.flatnonzero(gr.bounces > 0):
        b = min(gr.bounces[i], k)
        np.testing.assert_array_equal(gr.hit_patch[i, :b], orr.hit_patch[i, :b])
        assert np.abs(gr.hit_xyz[i, :b] - orr.hit_xyz[i, :b]).max() <= 1e-10


def test_bounce_loop_learned_seed_matches_triangle_oracle(c1, trained):
    sc, g, osc, o, d, mb = c1
    gr = g.trace(o, d, mb, mode=L.NEWTON_CONVERGED, seed=L.SEED_LEARNED)
    orr = osc.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    _compare(gr, orr)
    n_hits = gr.counter("seed_learned") + gr.counter("seed_fallback")
    tri = g.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    assert n_hits == tri.counter("newton_calls")
    assert gr.counter("seed_learned") > 0.9 * n_hits
    assert gr.counter("mlp_lanes") == n_hits and gr.counter("mlp_evals") > 0


def test_bad_network_changes_nothing(c1):
    """Large random weights: the seeds are garbage, the results still exactly the triangle path's."""
    sc, _, osc, o, d, mb = c1
    g = T.GpuScene(sc)
    net = T.Mlp(77)
    net.params = _params(78, 2.0)
    g.set_mlp(net)
    sel = slice(0, 4096, 3)
    gr = g.trace(o[sel], d[sel], mb, mode=L.NEWTON_CONVERGED, seed=L.SEED_LEARNED)
    orr = osc.trace(o[sel], d[sel], mb, mode=L.NEWTON_CONVERGED)
    _compare(gr, orr)
    assert gr.counter("seed_fallback") > 0


def test_learned_seed_argument_checks(c1):
    sc, _, _, o, d, mb = c1
    g = T.GpuScene(sc)
    with pytest.raises(L.BBRTError, match="bbrt_scene_set_mlp"):
        g.trace(o[:10], d[:10], mb, mode=L.NEWTON_CONVERGED, seed=L.SEED_LEARNED)
    g.set_mlp(T.Mlp(1))
    with pytest.raises(L.BBRTError, match="CONVERGED"):
        g.trace(o[:10], d[:10], mb, mode=L.NEWTON_REF, seed=L.SEED_LEARNED)
    net = T.Mlp(2)
    bad = net.params
    bad[M.OFF_W1 + 60] = 1.0   # a pad of W1
    with pytest.raises(L.BBRTError, match="padding"):
        net.params = bad


def test_batched_compact_learned_matches_triangle(trained):
    """The batched form at small scale: compact, device-built sub-scenes, learned seeds with fallback; the
    records equal the triangle-seed trace of the same handle, sub-scene 0 equals the oracle, and the per-ray
    learned/fallback counts add up to the counters."""
    net, _, _ = trained
    scenes, o, d, rb = S.batch_config(seed=8, n_scenes=5, patch_lo=80, patch_hi=600, rays_lo=300, rays_hi=1500)
    gb = T.GpuScene.batch(scenes, compact=True, device_build=True)
    gb.set_mlp(net)
    mb = 25
    gl = gb.trace(o, d, mb, mode=L.NEWTON_CONVERGED, seed=L.SEED_LEARNED, ray_begin=rb, seed_stats=True)
    tri = gb.trace(o, d, mb, mode=L.NEWTON_CONVERGED, ray_begin=rb)
    _compare(gl, tri)
    n_hits = gl.counter("seed_learned") + gl.counter("seed_fallback")
    assert n_hits == tri.counter("newton_calls") > 0
    assert gl.seed_stats[:, 0].sum() == gl.counter("seed_learned")
    assert gl.seed_stats[:, 1].sum() == gl.counter("seed_fallback")
    assert gl.counter("seed_learned") > 0
    sl = slice(rb[0], rb[1])
    part = SimpleNamespace(**{f: getattr(gl, f)[sl] for f in ("status", "err_kind", "bounces", "hit_patch", "hit_xyz")})
    _compare(part, O.OracleScene(scenes[0]).trace(o[sl], d[sl], mb, mode=L.NEWTON_CONVERGED))
