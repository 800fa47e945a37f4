"""ctypes wrapper of the CPU oracle (liboracle.so) and the reference-backed checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_i8 = C.POINTER(C.c_int8)
_u8 = C.POINTER(C.c_uint8)


class PatchSet(C.Structure):
    _fields_ = [("n_patches", C.c_int64), ("imax", _i32), ("jmax", _i32), ("u_off", _i64), ("v_off", _i64),
                ("f_off", _i64), ("unodes", _d), ("vnodes", _d), ("fields", _d)]


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_mt.restype = C.c_int
        L.orc_ray_box.restype = C.c_int
        L.orc_eval_field.restype = C.c_double
        L.orc_solve3x3.restype = C.c_int
        L.orc_bvh_build.restype = C.c_int64
        L.orc_bvh_visit.restype = C.c_int64
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.POINTER(PatchSet), C.c_int]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_diag.restype = C.c_double
        L.orc_scene_diag.argtypes = [C.c_void_p]
        L.orc_scene_n_nodes.restype = C.c_int64
        L.orc_scene_n_nodes.argtypes = [C.c_void_p]
        L.orc_scene_bvh.argtypes = [C.c_void_p, _d, _i32, _i32, _i32, _u8, _i32]
        L.orc_scene_patch_info.argtypes = [C.c_void_p, _d, _d, _d]
        L.orc_scene_coarse.argtypes = [C.c_void_p, C.c_int64, _d, _d, _d]
        L.orc_patch_intersect.restype = C.c_int
        L.orc_patch_intersect.argtypes = [C.c_void_p, C.c_int64, _d, _d, C.c_int, _d, C.POINTER(C.c_int)]
        L.orc_patch_intersect_uv.restype = C.c_int
        L.orc_patch_intersect_uv.argtypes = [C.c_void_p, C.c_int64, _d, _d, C.c_int, _d, C.POINTER(C.c_int)]
        L.orc_scene_intersect.restype = C.c_int
        L.orc_scene_intersect.argtypes = [C.c_void_p, _d, _d, C.c_int, C.c_int, _d, C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]
        L.orc_stage1_seed.restype = C.c_int
        L.orc_stage1_seed.argtypes = [C.c_void_p, C.c_int64, _d, _d, _d, C.POINTER(C.c_int)]
        L.orc_newton.restype = C.c_int
        L.orc_newton.argtypes = [C.c_void_p, C.c_int64, _d, _d, C.c_double, C.c_double, C.c_double, C.c_int, _d,
                                 C.POINTER(C.c_int)]
        L.orc_trace.argtypes = [C.c_void_p, _d, _d, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, _i8, _i8, _i32,
                                _i32, _i32, _d, _u8, _d, _i64]
        L.orc_max_threads.restype = C.c_int
        L.orc_bounce_map_launches.restype = None
        _LIB = L
    return _LIB


def ref_harness():
    """The reference-backed checker (oracle/_ref), or None when it is not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_harness.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_mt.restype = C.c_int
        R.ref_ray_box.restype = C.c_int
        R.ref_bvh_visit.restype = C.c_int64
        _REF = R
    return _REF


def _v(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def moller_trumbore(o, d, v0, v1, v2, eps=1e-9, fn="orc_mt"):
    """-> (hit, t, point, normal)."""
    L = lib() if fn == "orc_mt" else ref_harness()
    out = np.zeros(7)
    f = getattr(L, fn)
    h = f(ptr(_v(o), _d), ptr(_v(d), _d), ptr(_v(v0), _d), ptr(_v(v1), _d), ptr(_v(v2), _d), C.c_double(eps),
          ptr(out, _d))
    return bool(h), out[0], out[1:4], out[4:7]


def ray_box(o, inv, box, best_t, fn="orc_ray_box"):
    L = lib() if fn == "orc_ray_box" else ref_harness()
    tm = C.c_double(0.0)
    r = getattr(L, fn)(ptr(_v(o), _d), ptr(_v(inv), _d), ptr(_v(box), _d), C.c_double(best_t), C.byref(tm))
    return bool(r), tm.value


def bary_weights(nodes):
    nodes = _v(nodes)
    out = np.zeros_like(nodes)
    lib().orc_bary_weights(ptr(nodes, _d), C.c_int(nodes.size), ptr(out, _d))
    return out


def interp2d(xn, yn, F, x, y):
    xn, yn, F, x, y = _v(xn), _v(yn), _v(F), _v(x), _v(y)
    out = np.zeros((x.size, y.size))
    lib().orc_interp2d(ptr(xn, _d), C.c_int(xn.size), ptr(yn, _d), C.c_int(yn.size), ptr(F, _d), ptr(x, _d),
                       C.c_int(x.size), ptr(y, _d), C.c_int(y.size), ptr(out, _d))
    return out


def basis1d(nodes, x):
    nodes = _v(nodes)
    out = np.zeros_like(nodes)
    lib().orc_basis1d(ptr(nodes, _d), C.c_int(nodes.size), C.c_double(x), ptr(out, _d))
    return out


def eval_field(lu, lv, f, imax, jmax):
    return lib().orc_eval_field(ptr(_v(lu), _d), ptr(_v(lv), _d), ptr(_v(f), _d), C.c_int(imax), C.c_int(jmax))


def solve3x3(J, b):
    x = np.zeros(3)
    ok = lib().orc_solve3x3(ptr(_v(J).ravel(), _d), ptr(_v(b), _d), ptr(x, _d))
    return bool(ok), x


def bvh_visit(boxes, cents, obj_t, o, d, leaf_threshold=4, fn="orc_bvh_visit"):
    L = lib() if fn == "orc_bvh_visit" else ref_harness()
    boxes, cents, obj_t = _v(boxes), _v(cents), _v(obj_t)
    n = cents.shape[0]
    visits = np.zeros(4 * n + 16, np.int32)
    bo = C.c_int32(0); bt = C.c_double(0)
    nv = getattr(L, fn)(ptr(boxes, _d), ptr(cents, _d), C.c_int64(n), C.c_int(leaf_threshold), ptr(obj_t, _d),
                        ptr(_v(o), _d), ptr(_v(d), _d), ptr(visits, _i32), C.c_int64(visits.size), C.byref(bo),
                        C.byref(bt))
    return visits[:min(nv, visits.size)].copy(), bo.value, bt.value


def bvh_build(boxes, cents, leaf_threshold=4):
    boxes, cents = _v(boxes), _v(cents)
    n = cents.shape[0]
    nb = np.zeros((2 * n, 6)); r = np.zeros(2 * n, np.int32); s = np.zeros(2 * n, np.int32)
    c = np.zeros(2 * n, np.int32); a = np.zeros(2 * n, np.uint8); lo = np.zeros(n, np.int32)
    nn = lib().orc_bvh_build(ptr(boxes, _d), ptr(cents, _d), C.c_int64(n), C.c_int(leaf_threshold), ptr(nb, _d),
                             ptr(r, _i32), ptr(s, _i32), ptr(c, _i32), ptr(a, _u8), ptr(lo, _i32))
    return dict(box=nb[:nn], right=r[:nn], start=s[:nn], count=c[:nn], axis=a[:nn], leaf=lo)


@dataclass
class TraceResult:
    status: np.ndarray
    err_kind: np.ndarray
    bounces: np.ndarray
    err_patch: np.ndarray
    hit_patch: np.ndarray
    hit_xyz: np.ndarray
    iters: np.ndarray
    final_od: np.ndarray
    counters: np.ndarray


def patch_set(sc) -> PatchSet:
    ps = PatchSet()
    ps.n_patches = sc.n_patches
    ps.imax = ptr(sc.imax, _i32); ps.jmax = ptr(sc.jmax, _i32)
    ps.u_off = ptr(sc.u_off, _i64); ps.v_off = ptr(sc.v_off, _i64); ps.f_off = ptr(sc.f_off, _i64)
    ps.unodes = ptr(sc.unodes, _d); ps.vnodes = ptr(sc.vnodes, _d); ps.fields = ptr(sc.fields, _d)
    return ps


class OracleScene:
    """CPU oracle scene: patches prepared as SurfacePatch does, BVH as RayTracingDriver does."""

    def __init__(self, sc, threads: int = 0):
        self._sc = sc  # keeps the arrays alive
        self.n_patches = sc.n_patches
        self.h = lib().orc_scene_create(C.byref(patch_set(sc)), C.c_int(threads))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    @property
    def diag(self) -> float:
        return lib().orc_scene_diag(self.h)

    def bvh(self):
        n = lib().orc_scene_n_nodes(self.h)
        nb = np.zeros((n, 6)); r = np.zeros(n, np.int32); s = np.zeros(n, np.int32)
        c = np.zeros(n, np.int32); a = np.zeros(n, np.uint8); lo = np.zeros(self.n_patches, np.int32)
        lib().orc_scene_bvh(self.h, ptr(nb, _d), ptr(r, _i32), ptr(s, _i32), ptr(c, _i32), ptr(a, _u8), ptr(lo, _i32))
        return dict(box=nb, right=r, start=s, count=c, axis=a, leaf=lo)

    def patch_info(self):
        n = self.n_patches
        box = np.zeros((n, 6)); cent = np.zeros((n, 3)); ms = np.zeros(n)
        lib().orc_scene_patch_info(self.h, ptr(box, _d), ptr(cent, _d), ptr(ms, _d))
        return box, cent, ms

    def coarse(self, k: int):
        xyz = np.zeros((100 * 100, 3)); sag = np.zeros(99 * 99); cn = np.zeros((99 * 99, 3))
        lib().orc_scene_coarse(self.h, C.c_int64(k), ptr(xyz, _d), ptr(sag, _d), ptr(cn, _d))
        return xyz, sag, cn

    def patch_intersect(self, k, o, d, mode=0):
        out = np.zeros(7); it = C.c_int(0)
        st = lib().orc_patch_intersect(self.h, C.c_int64(k), ptr(_v(o), _d), ptr(_v(d), _d), C.c_int(mode),
                                       ptr(out, _d), C.byref(it))
        return st, out[0], out[1:4], out[4:7], it.value

    def patch_intersect_uv(self, k, o, d, mode=0):
        """(status, t, point, normal, iters, u, v) of SurfacePatch::intersect."""
        out = np.zeros(9); it = C.c_int(0)
        st = lib().orc_patch_intersect_uv(self.h, C.c_int64(k), ptr(_v(o), _d), ptr(_v(d), _d), C.c_int(mode),
                                          ptr(out, _d), C.byref(it))
        return st, out[0], out[1:4], out[4:7], it.value, out[7], out[8]

    def intersect(self, o, d, mode=0, brute=False):
        """BVH::intersect (or brute_force_intersect) over the patches, direction as given:
        (err_kind, err_patch, patch, t, point, normal)."""
        out = np.zeros(7); pt = C.c_int(-1); ep = C.c_int(-1)
        err = lib().orc_scene_intersect(self.h, ptr(_v(o), _d), ptr(_v(d), _d), C.c_int(mode), C.c_int(int(brute)),
                                        ptr(out, _d), C.byref(pt), C.byref(ep))
        return err, ep.value, pt.value, out[0], out[1:4], out[4:7]

    def stage1_seed(self, k, o, d):
        """Stage 1 of SurfacePatch::intersect alone: None on a miss, else (t0, u0, v0, i, j, second)."""
        out = np.zeros(3); idx = (C.c_int * 3)()
        h = lib().orc_stage1_seed(self.h, C.c_int64(k), ptr(_v(o), _d), ptr(_v(d), _d), ptr(out, _d), idx)
        if not h:
            return None
        return out[0], out[1], out[2], idx[0], idx[1], bool(idx[2])

    def newton(self, k, o, d, u0, v0, t0, mode=0):
        out = np.zeros(7); it = C.c_int(0)
        st = lib().orc_newton(self.h, C.c_int64(k), ptr(_v(o), _d), ptr(_v(d), _d), C.c_double(u0), C.c_double(v0),
                              C.c_double(t0), C.c_int(mode), ptr(out, _d), C.byref(it))
        return st, out[0], out[1:4], out[4:7], it.value

    def trace(self, origins, dirs, max_bounces, mode=1, record_k: Optional[int] = None, threads: int = 0):
        o = _v(origins).reshape(-1, 3); d = _v(dirs).reshape(-1, 3)
        n = o.shape[0]
        K = max_bounces if record_k is None else record_k
        K = max(K, 1)
        st = np.zeros(n, np.int8); ek = np.zeros(n, np.int8); b = np.zeros(n, np.int32); ep = np.zeros(n, np.int32)
        hp = np.zeros((n, K), np.int32); hx = np.zeros((n, K, 3)); it = np.zeros((n, K), np.uint8)
        fod = np.zeros((n, 6)); cnt = np.zeros(6, np.int64)
        lib().orc_trace(self.h, ptr(o, _d), ptr(d, _d), C.c_int64(n), C.c_int(mode), C.c_int(max_bounces),
                        C.c_int(K), C.c_int(threads), ptr(st, _i8), ptr(ek, _i8), ptr(b, _i32), ptr(ep, _i32),
                        ptr(hp, _i32), ptr(hx, _d), ptr(it, _u8), ptr(fod, _d), ptr(cnt, _i64))
        return TraceResult(st, ek, b, ep, hp, hx, it, fod, cnt)


def bounce_map_launches(osc, ids, n, theta_n, phi_n, cone_deg):
    """Launch rays of BounceMapDriver::run for scene patches `ids` (oracle restatement)."""
    ids = np.ascontiguousarray(ids, dtype=np.int64)
    m = ids.size * n * n * theta_n * phi_n
    o = np.zeros((m, 3)); d = np.zeros((m, 3))
    lib().orc_bounce_map_launches(C.c_void_p(osc.h), ptr(ids, _i64), C.c_int64(ids.size), C.c_int(n), C.c_int(theta_n),
                                  C.c_int(phi_n), C.c_double(cone_deg), ptr(o, _d), ptr(d, _d))
    return o, d


def max_threads() -> int:
    return lib().orc_max_threads()
