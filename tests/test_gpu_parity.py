"""GPU parity: the HIP bounce loop against the CPU FP64 oracle on identical inputs.

Integer outputs (status, error kind, bounce count, failing patch, per-bounce
patch sequence, Newton iteration counts) and the work counters that the
reference's control flow fixes (BVH node tests, stage-1 candidates, Newton
calls and iterations) must match exactly.  Hit points and final ray states
are compared bit-for-bit: both sides evaluate the same FP64 expressions in
the same order without contraction, so the tolerance is zero.
"""
import numpy as np
import pytest

from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from billardballraytracing3d_amd import _lib as L

pytestmark = pytest.mark.gpu

_CACHE = {}


def _scene(name):
    if name not in _CACHE:
        if name == "two_spheres":
            sc, o, d, mb = S.two_spheres_plane_source()
        elif name == "sinai_sample":
            sc, o, d, mb = S.sinai_cone(n_rays=2048)
        else:
            raise KeyError(name)
        _CACHE[name] = (sc, o, d, mb)
    return _CACHE[name]


def _oracle(name, mode, sc, o, d, mb):
    from oracle import oracle as O
    key = (name, "oracle", mode)
    if key not in _CACHE:
        osc = O.OracleScene(sc)
        _CACHE[key] = osc.trace(o, d, mb, mode=mode)
    return _CACHE[key]


def _gpu_scene(name, sc):
    key = (name, "gpu")
    if key not in _CACHE:
        _CACHE[key] = T.GpuScene(sc)
    return _CACHE[key]


def assert_same(g, r, label):
    np.testing.assert_array_equal(g.status, r.status, err_msg=label + " status")
    np.testing.assert_array_equal(g.err_kind, r.err_kind, err_msg=label + " err_kind")
    np.testing.assert_array_equal(g.bounces, r.bounces, err_msg=label + " bounces")
    np.testing.assert_array_equal(g.err_patch, r.err_patch, err_msg=label + " err_patch")
    np.testing.assert_array_equal(g.hit_patch, r.hit_patch, err_msg=label + " hit_patch")
    np.testing.assert_array_equal(g.iters, r.iters, err_msg=label + " newton iters")
    np.testing.assert_array_equal(g.hit_xyz, r.hit_xyz, err_msg=label + " hit_xyz (bit-exact)")
    np.testing.assert_array_equal(g.final_od, r.final_od, err_msg=label + " final origin/direction")


@pytest.mark.parametrize("name", ["two_spheres", "sinai_sample"])
@pytest.mark.parametrize("mode", [L.NEWTON_REF, L.NEWTON_CONVERGED])
def test_trace_matches_oracle(name, mode):
    sc, o, d, mb = _scene(name)
    print(f"[parity] scene={name} mode={mode} rays={o.shape[0]} patches={sc.n_patches}", flush=True)
    r = _oracle(name, mode, sc, o, d, mb)
    g = _gpu_scene(name, sc).trace(o, d, mb, mode=mode)
    assert_same(g, r, f"{name}/mode{mode}")
    # counters the reference's control flow fixes: [calls, iters, candidates, mt, box tests, polish fail]
    assert g.counter("newton_calls") == r.counters[0]
    assert g.counter("newton_iters") == r.counters[1]
    assert g.counter("candidates") == r.counters[2]
    assert g.counter("polish_fail") == r.counters[5]
    # the per-bounce schedule walks the binary BVH node by node, as the reference does: its node tests are
    # the reference's; the persistent kernel's 4-wide walk visits fewer, wider nodes (same leaves, same order)
    w = _gpu_scene(name, sc).trace(o, d, mb, mode=mode, flags=L.FLAG_WAVEFRONT)
    assert_same(w, r, f"{name}/mode{mode} wavefront")
    assert w.counter("bvh_node_tests") == r.counters[4]
    assert 0 < g.counter("bvh_node_tests") < w.counter("bvh_node_tests")
    assert g.counter("bounces") == int(r.bounces.sum())
    assert g.counter("rays") == o.shape[0]


@pytest.mark.parametrize("name", ["two_spheres", "sinai_sample"])
def test_wavefront_schedule_matches_persistent(name):
    sc, o, d, mb = _scene(name)
    gs = _gpu_scene(name, sc)
    print(f"[schedule] scene={name}", flush=True)
    a = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    b = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED, flags=L.FLAG_WAVEFRONT)
    assert_same(a, b, name + " persistent vs wavefront")
    for k in ("bounces", "candidates", "cull_box_tests", "mt_tests", "newton_iters"):
        assert a.counter(k) == b.counter(k), k
    assert a.counter("bvh_node_tests") < b.counter("bvh_node_tests")   # 4-wide walk vs the binary one


@pytest.mark.parametrize("name", ["two_spheres", "sinai_sample"])
def test_wave_cooperative_stage1_matches_per_lane(name):
    """Stage 1 run by the whole wave for every candidate, or by each lane alone, gives the default run's
    records and work counters."""
    sc, o, d, mb = _scene(name)
    gs = _gpu_scene(name, sc)
    print(f"[stage1 wave/lane] scene={name}", flush=True)
    a = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    for fl in (L.FLAG_STAGE1_WAVE, L.FLAG_STAGE1_LANE, L.FLAG_STAGE1_WAVE | L.FLAG_NO_CULL):
        b = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED, flags=fl)
        assert_same(a, b, f"{name} flags {fl}")
        keys = ("bounces", "candidates", "newton_calls", "newton_iters") + \
            (("cull_box_tests", "mt_tests", "bvh_node_tests") if not fl & L.FLAG_NO_CULL else ())
        for k in keys:
            assert a.counter(k) == b.counter(k), (fl, k)


@pytest.mark.parametrize("name", ["two_spheres", "sinai_sample"])
def test_culling_equals_brute_force(name):
    sc, o, d, mb = _scene(name)
    gs = _gpu_scene(name, sc)
    print(f"[cull] scene={name}", flush=True)
    a = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    b = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED, flags=L.FLAG_NO_CULL)
    assert_same(a, b, name + " culled vs brute force")
    assert b.counter("mt_tests") == 6 * 99 * 99 * b.counter("candidates")
    assert a.counter("mt_tests") < b.counter("mt_tests")


def test_edge_cases_match_oracle():
    """Zero directions, rays starting inside a sphere (back face), axis-parallel
    rays, max_bounces 0 and 1, a ragged ray count, records shorter than the
    bounce cap, and an empty batch."""
    from oracle import oracle as O
    sc, o, d, mb = _scene("two_spheres")
    rng = np.random.default_rng(3)
    n = 333
    oo = rng.uniform(-3, 3, (n, 3))
    dd = rng.normal(size=(n, 3))
    dd[::11] = 0.0                                    # zero-length: Error, zero_dir
    oo[1::13] = [-1.5, 0.1, 0.2]                      # inside the left sphere: back-face hits
    dd[2::17] = [[1.0, 0.0, 0.0]]                     # axis-parallel: 0 * inf in the slab test
    gs = _gpu_scene("two_spheres", sc)
    osc = O.OracleScene(sc)
    for max_b, rec in [(0, 1), (1, 1), (20, 20), (20, 5)]:
        print(f"[edge] max_bounces={max_b} record_k={rec}", flush=True)
        g = gs.trace(oo, dd, max_b, mode=L.NEWTON_REF, record_k=rec)
        r = osc.trace(oo, dd, max_b, mode=0, record_k=rec)
        assert_same(g, r, f"edge max_b={max_b}")
    assert (g.err_kind == L.ERR_ZERO_DIR).sum() == int(np.all(dd == 0.0, axis=1).sum()) > 0
    assert (g.err_kind == L.ERR_BACKFACE).sum() > 0
    e = gs.trace(np.zeros((0, 3)), np.zeros((0, 3)), 20)
    assert e.status.size == 0 and e.counter("rays") == 0


def test_far_origins_match_oracle():
    """Rays starting beyond the culling limit (8 x the scene scale): the first bounce walks the BVH with FP64 slot
    tests and runs brute-force stage 1; later bounces start inside the scene and use the FP32 walk."""
    from oracle import oracle as O
    sc, o, d, mb = _scene("two_spheres")
    rng = np.random.default_rng(5)
    n = 96
    target = rng.uniform(-2.0, 2.0, (n, 3)) * [1.0, 0.5, 0.5]
    dirn = rng.normal(size=(n, 3))
    dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
    oo = target - 60.0 * dirn                          # 60 units out, aimed at the spheres
    gs = _gpu_scene("two_spheres", sc)
    print(f"[far] rays={n}", flush=True)
    for mode in (L.NEWTON_REF, L.NEWTON_CONVERGED):
        g = gs.trace(oo, dirn, mb, mode=mode)
        r = O.OracleScene(sc).trace(oo, dirn, mb, mode=mode)
        assert_same(g, r, f"far origins mode{mode}")
        assert g.counter("bounces") > n // 4
        assert g.counter("candidates") == r.counters[2]


def test_hit_counts_on_device():
    """PlaneSourceDriver semantics (PlaneSourceDriver.hpp:96-108) accumulated on the GPU."""
    sc, o, d, mb = _scene("two_spheres")
    gs = _gpu_scene("two_spheres", sc)
    for direct in (False, True):
        g = gs.trace(o, d, mb, mode=L.NEWTON_REF, hit_count=True, flags=L.FLAG_COUNT_DIRECT_HIT if direct else 0)
        want = np.zeros(sc.n_patches, np.int64)
        for i in range(o.shape[0]):
            for k in range(0 if direct else 1, g.bounces[i]):
                want[g.hit_patch[i, k]] += 1
        np.testing.assert_array_equal(g.hit_count, want)


@pytest.mark.parametrize("seed", [21, 22, 23])
def test_random_scenes_match_oracle(seed):
    """Random multi-obstacle scenes (1-8 spheres/ellipsoids, order 2 or 4, uneven patch counts) with 30 degree
    cones of rays: every record bit-identical to the oracle, in the reference's Newton mode."""
    from oracle import oracle as O
    scenes, _, _, _ = S.batch_config(seed=seed, n_scenes=1, patch_lo=80, patch_hi=900)
    sc = scenes[0]
    # rays from a sphere of radius 10 aimed at points inside the obstacles, so most hit and many bounce
    rng = np.random.default_rng(seed)
    n = 1500
    ob = [sc.obstacles[k] for k in rng.integers(0, len(sc.obstacles), n)]
    target = np.array([b.center + rng.uniform(-0.8, 0.8, 3) * np.asarray(b.axes) for b in ob])
    o = rng.normal(size=(n, 3))
    o = 10.0 * o / np.linalg.norm(o, axis=1, keepdims=True)
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    print(f"[random] seed={seed} patches={sc.n_patches} rays={n}", flush=True)
    g = T.GpuScene(sc).trace(o, d, 50, mode=L.NEWTON_REF)
    r = O.OracleScene(sc).trace(o, d, 50, mode=L.NEWTON_REF)
    assert_same(g, r, f"random seed {seed}")
    assert g.counter("candidates") == r.counters[2] and g.counter("newton_iters") == r.counters[1]
    assert g.counter("bounces") > n // 2
