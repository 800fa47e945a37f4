"""Stage 1 under load.  Rays that run along the flat top of a thin ellipsoid (semi-axes 1, 1, 1e-4) inside its
cells' padded boxes select about a dozen coarse cells per candidate patch, so with 24 or more leaf lanes a round a
wave's breadth-first walk (the default) needs several passes of 64 entries at its cell level; rays that graze a
sphere of eight large order-4 patches mix misses at every level with deep walks.  The records and work counters
must equal those of every lane walking its own pyramid (BBRT_FLAG_STAGE1_LANE) and of the whole wave walking one
lane's pyramid at a time (BBRT_FLAG_STAGE1_WAVE), and a sample must equal the FP64 oracle (brute-force stage 1)
exactly.
"""
import numpy as np
import pytest

from billardballraytracing3d_amd import _lib as L
from billardballraytracing3d_amd import scenes as S
from billardballraytracing3d_amd import tracer as T
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _grazing_sphere(n, seed):
    """Rays from x = -3 along +x (tilted by up to 1e-3) at distances 0.98-1.005 from the x axis: they pass the
    unit sphere's surface tangentially or skim inside it."""
    rng = np.random.default_rng(seed)
    h = 1.0 + rng.uniform(-0.02, 0.005, n)
    ang = rng.uniform(0.0, 2.0 * np.pi, n)
    o = np.stack([np.full(n, -3.0), h * np.cos(ang), h * np.sin(ang)], 1)
    d = np.stack([np.ones(n), rng.uniform(-1e-3, 1e-3, n), rng.uniform(-1e-3, 1e-3, n)], 1)
    return o, d


def _grazing_flat(n, seed, c):
    """Rays from x = -3 along +x (tilted by up to 0.05 in y and 5e-5 in z) at the height c of the flat top of an
    ellipsoid with semi-axes (1, 1, c): they run along the top face inside its cells' padded boxes."""
    rng = np.random.default_rng(seed)
    o = np.stack([np.full(n, -3.0), rng.uniform(-0.3, 0.3, n), np.full(n, c)], 1)
    d = np.stack([np.ones(n), rng.uniform(-0.05, 0.05, n), rng.uniform(-5e-5, 5e-5, n)], 1)
    return o, d


CASES = {
    # (obstacle, rays, fewest coarse cells tested per stage-1 candidate)
    "grazing_sphere": (lambda: S.Obstacle("sphere", np.array([0.0, 0.0, 0.0]), np.array([1.0, 1.0, 1.0]), "oct", 1,
                                          4, rot_seed=7),
                       lambda: _grazing_sphere(1 << 15, 11), 0.5),
    "flat_ellipsoid": (lambda: S.Obstacle("ellipsoid", np.array([0.0, 0.0, 0.0]), np.array([1.0, 1.0, 1e-4]), "oct", 1,
                                          4, rot_seed=7, rotate=False),
                       lambda: _grazing_flat(1 << 14, 3, 1e-4), 8.0),
}


def _same(a, b, label):
    for f in ("status", "err_kind", "bounces", "err_patch", "hit_patch", "iters", "hit_xyz", "final_od"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f"{label} {f}")


@pytest.mark.parametrize("name", list(CASES))
def test_breadth_first_stage1_under_load(name):
    make_ob, make_rays, min_cells = CASES[name]
    sc = S.build_scene([make_ob()])
    o, d = make_rays()
    mb = 6
    gs = T.GpuScene(sc)
    print(f"[grazing] {name}: {sc.n_patches} patches, {o.shape[0]} rays", flush=True)
    a = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED)
    cand = a.counter("candidates")
    print(f"[grazing] {name}: candidates {cand}, cull tests {a.counter('cull_box_tests')}, MT tests "
          f"{a.counter('mt_tests')}, bounces {a.counter('bounces')}", flush=True)
    # cells per candidate: with 24 or more leaf lanes a round, the flat case's cell level takes several passes
    assert a.counter("mt_tests") >= 6 * min_cells * cand
    assert a.counter("bounces") > 1000
    for fl in (L.FLAG_STAGE1_LANE, L.FLAG_STAGE1_WAVE):
        b = gs.trace(o, d, mb, mode=L.NEWTON_CONVERGED, flags=fl)
        _same(a, b, f"{name} flags {fl}")
        for k in ("bounces", "candidates", "cull_box_tests", "mt_tests", "bvh_node_tests", "newton_calls",
                  "newton_iters"):
            assert a.counter(k) == b.counter(k), (name, fl, k)
    # a sample against the FP64 oracle (brute-force stage 1)
    sel = slice(0, o.shape[0], 64)
    r = O.OracleScene(sc).trace(o[sel], d[sel], mb, mode=L.NEWTON_CONVERGED)
    for f in ("status", "err_kind", "bounces", "err_patch", "hit_patch", "iters", "hit_xyz", "final_od"):
        np.testing.assert_array_equal(getattr(a, f)[sel], getattr(r, f), err_msg=f"{name} oracle {f}")
    assert int(r.bounces.sum()) > 20
