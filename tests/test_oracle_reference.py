"""The CPU oracle against the reference's own compiled code and its known-answer test.

Moller-Trumbore, ray_box and the BVH (build + traversal order, pruning, ties)
are pinned bit-for-bit against fixtures written from the reference headers
(tests/golden/make_golden.py).  When oracle/_ref is built (a container with the
reference checkout), the same checks also run live against it on fresh
random inputs, and the reference's own tests/Triangle_Intersect.cpp binary is
run.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF_TEST = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "triangle_intersect_test")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def same_float(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64)) or np.array_equal(a, b, equal_nan=True)


def test_moller_trumbore_matches_reference_fixture():
    g = load("mt_cases.npz")
    rays, tris, hit, out = g["rays"], g["tris"], g["hit"], g["out"]
    assert hit.sum() > 100 and (hit == 0).sum() > 100
    for k in range(rays.shape[0]):
        h, t, p, n = O.moller_trumbore(rays[k, :3], rays[k, 3:], tris[k, :3], tris[k, 3:6], tris[k, 6:])
        assert h == bool(hit[k]), k
        if h:
            assert same_float(np.concatenate([[t], p, n]), out[k]), k


def test_ray_box_matches_reference_fixture():
    g = load("raybox_cases.npz")
    assert g["hit"].sum() > 100
    for k in range(g["o"].shape[0]):
        r, tm = O.ray_box(g["o"][k], g["inv"][k], g["box"][k], g["best"][k])
        assert r == bool(g["hit"][k]), k
        if r:
            assert same_float(tm, g["tmin"][k]), k


def test_bvh_visit_order_matches_reference_fixture():
    g = load("bvh_cases.npz")
    off, voff = g["obj_off"], g["visit_off"]
    nv = np.diff(voff)
    assert (nv > 0).sum() >= 70 and nv.max() >= 10   # the rays reach leaves: the visit order is exercised
    for c in range(g["n"].shape[0]):
        sl = slice(off[c], off[c + 1])
        vis, best, bt = O.bvh_visit(g["boxes"][sl], g["cents"][sl], g["obj_t"][sl], g["o"][c], g["d"][c],
                                    int(g["leaf_threshold"][c]))
        np.testing.assert_array_equal(vis, g["visits"][voff[c]:voff[c + 1]], err_msg=f"case {c}")
        assert best == g["best"][c]
        assert same_float(bt, g["best_t"][c])


def test_known_answer_scene():
    """tests/Triangle_Intersect.cpp:26-31: t = 1, 1, miss, 5, BVH == brute force."""
    g = load("kat.npz")
    np.testing.assert_array_equal(g["bvh_t"], [1.0, 1.0, np.nan, 5.0])
    tris, rays = g["tris"], g["rays"]
    boxes = np.concatenate([tris.reshape(4, 3, 3).min(1), tris.reshape(4, 3, 3).max(1)], 1)
    cents = tris.reshape(4, 3, 3).sum(1) / 3.0
    for r in range(4):
        ts = np.array([O.moller_trumbore(rays[r, :3], rays[r, 3:], t[:3], t[3:6], t[6:])[1]
                       if O.moller_trumbore(rays[r, :3], rays[r, 3:], t[:3], t[3:6], t[6:])[0] else np.nan
                       for t in tris])
        vis, best, bt = O.bvh_visit(boxes, cents, ts, rays[r, :3], rays[r, 3:], leaf_threshold=1)
        assert same_float(bt, g["bvh_t"][r]) and same_float(bt, g["brute_t"][r])


@pytest.mark.skipif(not os.path.exists(REF_TEST), reason="reference KAT binary not built (no reference checkout)")
def test_reference_kat_binary_passes():
    r = subprocess.run([REF_TEST], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "All tests passed" in r.stdout


@pytest.fixture
def live_reference():
    """The reference-backed checker, loaded only when a test asks for it (never at collection, so a
    `-m gpu` run does not load it)."""
    if O.ref_harness() is None:
        pytest.skip("oracle/_ref not built (no reference checkout)")


def test_oracle_matches_live_reference_random(live_reference):
    rng = np.random.default_rng(7)
    for _ in range(400):
        v = rng.uniform(-2, 2, (3, 3)); o = rng.uniform(-4, 4, 3)
        d = (rng.dirichlet([1, 1, 1]) @ v - o) if rng.uniform() < 0.7 else rng.normal(size=3)
        a = O.moller_trumbore(o, d, v[0], v[1], v[2])
        b = O.moller_trumbore(o, d, v[0], v[1], v[2], fn="ref_mt")
        assert a[0] == b[0]
        if a[0]:
            assert same_float(np.concatenate([[a[1]], a[2], a[3]]), np.concatenate([[b[1]], b[2], b[3]]))
    for trial in range(100):
        n = int(rng.integers(1, 40))
        c = rng.uniform(-5, 5, (n, 3)); h = rng.uniform(0.05, 1.5, (n, 3))
        boxes = np.concatenate([c - h, c + h], 1)
        t = rng.uniform(0.5, 20, n); t[rng.uniform(size=n) < 0.4] = np.nan
        o = rng.uniform(-8, 8, 3); d = rng.normal(size=3); d /= np.linalg.norm(d)
        lt = int(rng.integers(1, 6))
        a = O.bvh_visit(boxes, c, t, o, d, lt)
        b = O.bvh_visit(boxes, c, t, o, d, lt, fn="ref_bvh_visit")
        np.testing.assert_array_equal(a[0], b[0])
        assert a[1] == b[1]


def test_bvh_equals_brute_force_on_triangle_soups():
    """The reference's own oracle pattern (BruteForce.hpp:7-15): nearest hit from the BVH == scan."""
    rng = np.random.default_rng(11)
    for _ in range(60):
        n = int(rng.integers(1, 60))
        tris = rng.uniform(-5, 5, (n, 1, 3)) + rng.uniform(-1, 1, (n, 3, 3))
        boxes = np.concatenate([tris.min(1), tris.max(1)], 1)
        cents = tris.sum(1) / 3.0
        for _ in range(5):
            o = rng.uniform(-8, 8, 3)
            d = tris[int(rng.integers(0, n))].mean(0) - o
            ts = []
            for t in tris:
                h = O.moller_trumbore(o, d, t[0], t[1], t[2])
                ts.append(h[1] if h[0] else np.nan)
            ts = np.array(ts)
            vis, best, bt = O.bvh_visit(boxes, cents, ts, o, d, leaf_threshold=int(rng.integers(1, 5)))
            if np.all(np.isnan(ts)):
                assert best == -1
            else:
                assert bt == np.nanmin(ts)
