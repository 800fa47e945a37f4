"""The 4-wide BVH walk tests node boxes in FP32 (widened by the scene's delta, rounded outward) and leaves in FP64.
Its records equal the binary FP64 walk's only if, for every ray whose origin lies within the culling limit
(|o|_inf <= 8 S), an FP64 pass implies an FP32 pass with an entry t no larger than the FP64 one (plus the
1e-6 relative prune slack).  This checks that implication on adversarial rays, with numpy's IEEE float32
arithmetic standing in for the device's (no FMA contraction, correctly rounded division)."""
import numpy as np

F32 = np.float32
PRUNE_SLACK = 1.000001


def slab64(o, inv, lo, hi):
    """ray_box's best_t-independent part (BVH.hpp:14-37): pass flag and tmin, std::max/min NaN semantics."""
    tmin, tmax = -np.inf, np.inf
    for a in range(3):
        t1 = (lo[a] - o[a]) * inv[a]
        t2 = (hi[a] - o[a]) * inv[a]
        if inv[a] < 0.0:
            t1, t2 = t2, t1
        tmin = t1 if tmin < t1 else tmin
        tmax = t2 if t2 < tmax else tmax
        if tmax < tmin:
            return False, tmin
    return bool(tmax >= (0.0 if tmin < 0.0 else tmin)), tmin


def round_down(v):
    f = F32(v)
    return np.nextafter(f, F32(-np.inf)) if float(f) > v else f


def round_up(v):
    f = F32(v)
    return np.nextafter(f, F32(np.inf)) if float(f) < v else f


def slab32(o32, i32, lo32, hi32):
    tmin, tmax = F32(-np.inf), F32(np.inf)
    for a in range(3):
        t1 = F32(F32(lo32[a] - o32[a]) * i32[a])
        t2 = F32(F32(hi32[a] - o32[a]) * i32[a])
        if i32[a] < 0:
            t1, t2 = t2, t1
        tmin = t1 if tmin < t1 else tmin
        tmax = t2 if t2 < tmax else tmax
    return bool(tmax >= tmin and tmax >= 0), tmin


def test_fp32_widened_boxes_contain_the_fp64_test():
    rng = np.random.default_rng(11)
    S = 5.0                                         # scene scale: max(1, |coordinate|)
    delta = 1e-5 * S                                # the smallest widening the scene build uses
    checked = passed64 = 0
    with np.errstate(all="ignore"):
        for trial in range(40000):
            c = rng.uniform(-S, S, 3)
            ext = 10.0 ** rng.uniform(-3, 0.3, 3)
            lo, hi = np.minimum(c - ext, S), np.maximum(np.minimum(c + ext, S), c - ext)
            # a target on the box's boundary (face, edge or corner), nudged by a few ulps either way
            t = rng.uniform(lo, hi)
            for a in np.flatnonzero(rng.random(3) < 0.6):
                t[a] = lo[a] if rng.random() < 0.5 else hi[a]
            t = t * (1.0 + rng.choice([-1, 0, 1], 3) * 2.0 ** -50)
            o = rng.uniform(-8 * S, 8 * S, 3)
            if rng.random() < 0.2:                      # origins inside or on the box too
                o = rng.uniform(lo, hi)
            d = t - o
            if rng.random() < 0.15:                     # axis-parallel and nearly parallel directions
                d[rng.integers(3)] = 0.0 if rng.random() < 0.5 else 1e-300
            n = np.linalg.norm(d)
            if not n > 0:
                continue
            d = d / n
            inv = 1.0 / d
            ok64, tmin64 = slab64(o, inv, lo, hi)
            checked += 1
            if not ok64:
                continue
            passed64 += 1
            lo32 = [round_down(lo[a] - delta) for a in range(3)]
            hi32 = [round_up(hi[a] + delta) for a in range(3)]
            o32 = [F32(o[a]) for a in range(3)]
            i32 = [F32(1.0) / F32(d[a]) for a in range(3)]
            ok32, tmin32 = slab32(o32, i32, lo32, hi32)
            assert ok32, (trial, o, d, lo, hi)
            if tmin64 >= 0.0:
                assert float(tmin32) <= tmin64 * PRUNE_SLACK or np.isnan(tmin32), (trial, tmin32, tmin64)
    assert checked > 35000 and passed64 > 5000, (checked, passed64)


def test_the_check_catches_too_small_a_widening():
    """With no widening, FP32 rounding of the origin loses grazing hits: the implication above is not vacuous."""
    rng = np.random.default_rng(12)
    misses = 0
    with np.errstate(all="ignore"):
        for _ in range(4000):
            lo = rng.uniform(-5, 4, 3)
            hi = lo + rng.uniform(1e-3, 1.0, 3)
            t = np.where(rng.random(3) < 0.5, lo, hi)   # a corner: the ray grazes three faces at once
            o = rng.uniform(-40, 40, 3)
            d = (t - o) / np.linalg.norm(t - o)
            ok64, _ = slab64(o, 1.0 / d, lo, hi)
            if not ok64:
                continue
            ok32, _ = slab32([F32(x) for x in o], [F32(1.0) / F32(x) for x in d], [F32(x) for x in lo],
                             [F32(x) for x in hi])
            misses += not ok32
    assert misses > 0
