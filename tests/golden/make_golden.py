"""Writes the golden fixtures of tests/golden/ from the reference's own compiled code.

Runs in a container that has the reference checkout: `make -C oracle` builds
oracle/_ref/libref_harness.so from the reference headers in place (BVH.hpp,
MollerTrumbore.hpp, Triangle.hpp, BruteForce.hpp; SurfacePatch.hpp is not
buildable here: it needs <mkl.h>).  The fixtures are inputs and the
reference's outputs only:

  mt_cases.npz      Moller-Trumbore (MollerTrumbore.hpp:32-64): rays, triangles, hit flag, t, point, normal
  raybox_cases.npz  ray_box (BVH.hpp:14-37) incl. zero direction components (0 * inf = NaN) and best_t prunes
  bvh_cases.npz     BVH(objects, threshold).intersect (BVH.hpp:65-189) over probe objects: visit order, winner;
                    7 of 8 rays aimed near an object's centroid, so most walks reach several leaves
  kat.npz           the known-answer scene of tests/Triangle_Intersect.cpp:14-31 (BVH and brute force t)

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def mt_cases(rng):
    n = 600
    rays, tris = [], []
    for k in range(n):
        v = rng.uniform(-2, 2, (3, 3))
        o = rng.uniform(-4, 4, 3)
        kind = k % 6
        if kind == 0:      # aimed at an interior point
            w = rng.dirichlet([1, 1, 1])
            tgt = w @ v
            d = tgt - o
        elif kind == 1:    # aimed at an edge / vertex exactly
            a, b = rng.choice(3, 2, replace=False)
            s = [0.0, 0.5, 1.0][k % 3]
            tgt = v[a] * (1 - s) + v[b] * s
            d = tgt - o
        elif kind == 2:    # nearly parallel to the plane
            nrm = np.cross(v[1] - v[0], v[2] - v[0])
            d = rng.normal(size=3)
            d -= nrm * (d @ nrm) / (nrm @ nrm)
            d += nrm * rng.uniform(-1e-9, 1e-9)
        elif kind == 3:    # axis-aligned directions, some components exactly zero
            d = np.zeros(3)
            d[k % 3] = rng.choice([-1.0, 1.0])
        elif kind == 4:    # behind the origin (t < 0) or tiny t
            w = rng.dirichlet([1, 1, 1])
            tgt = w @ v
            d = o - tgt if k % 2 else (tgt - o) * 1e9
        else:              # random
            d = rng.normal(size=3)
        rays.append(np.concatenate([o, d]))
        tris.append(v.ravel())
    rays = np.array(rays); tris = np.array(tris)
    hit = np.zeros(n, np.int8); out = np.full((n, 7), np.nan)
    for k in range(n):
        h, t, p, nn = O.moller_trumbore(rays[k, :3], rays[k, 3:], tris[k, :3], tris[k, 3:6], tris[k, 6:], 1e-9,
                                        fn="ref_mt")
        hit[k] = h
        if h:
            out[k] = np.concatenate([[t], p, nn])
    return dict(rays=rays, tris=tris, hit=hit, out=out)


def raybox_cases(rng):
    n = 500
    o = rng.uniform(-3, 3, (n, 3))
    d = rng.normal(size=(n, 3))
    d[rng.uniform(size=(n, 3)) < 0.2] = 0.0
    d[np.all(d == 0, axis=1), 2] = 1.0
    lo = rng.uniform(-2, 1, (n, 3)); hi = lo + rng.uniform(0, 2, (n, 3))
    box = np.concatenate([lo, hi], 1)
    aim = rng.uniform(size=n) < 0.6          # aimed at a point inside the box
    tgt = lo + rng.uniform(size=(n, 3)) * (hi - lo)
    d[aim] = tgt[aim] - o[aim]
    # origins exactly on a box face in some cases (0 * inf = NaN in the slab test)
    on = rng.uniform(size=n) < 0.15
    o[on, 0] = lo[on, 0]
    best = np.where(rng.uniform(size=n) < 0.5, np.inf, rng.uniform(0, 5, n))
    inv = np.empty_like(d)
    with np.errstate(divide="ignore"):
        inv[:] = 1.0 / d
    res = np.zeros(n, np.int8); tmin = np.full(n, np.nan)
    for k in range(n):
        r, tm = O.ray_box(o[k], inv[k], box[k], best[k], fn="ref_ray_box")
        res[k] = r
        tmin[k] = tm
    return dict(o=o, inv=inv, box=box, best=best, hit=res, tmin=tmin)


def bvh_cases(rng):
    cases = []
    for trial in range(80):
        n = int(rng.integers(1, 48))
        c = rng.uniform(-5, 5, (n, 3))
        if trial % 5 == 0:
            c = np.round(c)          # tied centroids
        h = rng.uniform(0.05, 1.5, (n, 3))
        boxes = np.concatenate([c - h, c + h], 1)
        t = rng.uniform(0.5, 20, n)
        t[rng.uniform(size=n) < 0.4] = np.nan
        if trial % 7 == 0:
            t[:] = 5.0               # tied hits
        o = rng.uniform(-8, 8, 3)
        if trial % 8 == 7:           # a random direction
            d = rng.normal(size=3)
        else:                        # aimed at a point near one object's centroid, so the walk reaches leaves
            tgt = c[int(rng.integers(0, n))] + rng.uniform(-0.5, 0.5, 3) * h[0]
            d = tgt - o
        if trial % 3 == 0:           # a zero direction component (0 * inf = NaN in the slab tests)
            k = int(rng.integers(0, 3))
            if trial % 8 != 7:
                o[k] = tgt[k]        # keeps the aimed ray on its target
            d[k] = 0.0
        d /= np.linalg.norm(d)
        lt = int(rng.integers(1, 6))
        vis, best, bt = O.bvh_visit(boxes, c, t, o, d, lt, fn="ref_bvh_visit")
        cases.append((n, lt, boxes, c, t, o, d, vis, best, bt))
    N = np.array([q[0] for q in cases]); LT = np.array([q[1] for q in cases])
    off = np.concatenate([[0], np.cumsum(N)])
    nv = np.array([len(q[7]) for q in cases]); voff = np.concatenate([[0], np.cumsum(nv)])
    return dict(n=N, leaf_threshold=LT, obj_off=off, boxes=np.concatenate([q[2] for q in cases]),
                cents=np.concatenate([q[3] for q in cases]), obj_t=np.concatenate([q[4] for q in cases]),
                o=np.array([q[5] for q in cases]), d=np.array([q[6] for q in cases]), visit_off=voff,
                visits=np.concatenate([q[7] for q in cases]).astype(np.int32),
                best=np.array([q[8] for q in cases], np.int32), best_t=np.array([q[9] for q in cases]))


def kat():
    # tests/Triangle_Intersect.cpp:14-31
    tris = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [10, 0, 0, 11, 0, 0, 10, 1, 0],
                     [0, 0, 5, 1, 0, 5, 0, 1, 5], [-5, 0, 0, -4, 0, 0, -5, 1, 0]], dtype=np.float64)
    rays = np.array([[0.2, 0.2, -1, 0, 0, 1], [10.2, 0.2, -1, 0, 0, 1], [100, 100, -1, 0, 0, 1],
                     [0.2, 0.2, 10, 0, 0, -1]], dtype=np.float64)
    import ctypes as C
    R = O.ref_harness()
    out = np.zeros(8)
    R.ref_triangle_kat(tris.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(4),
                       rays.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(4), C.c_int(1),
                       out.ctypes.data_as(C.POINTER(C.c_double)))
    return dict(tris=tris, rays=rays, bvh_t=out[:4], brute_t=out[4:])


def main():
    if O.ref_harness() is None:
        raise SystemExit("oracle/_ref/libref_harness.so is not built (make -C oracle with the reference present)")
    rng = np.random.default_rng(20261203)
    np.savez_compressed(os.path.join(HERE, "mt_cases.npz"), **mt_cases(rng))
    np.savez_compressed(os.path.join(HERE, "raybox_cases.npz"), **raybox_cases(rng))
    np.savez_compressed(os.path.join(HERE, "bvh_cases.npz"), **bvh_cases(rng))
    np.savez_compressed(os.path.join(HERE, "kat.npz"), **kat())
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
