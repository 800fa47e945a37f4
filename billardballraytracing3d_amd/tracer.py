"""Python front end of the HIP bounce loop (libbbrt_hip.so).

``GpuScene(scene)`` uploads a patch set; ``trace`` runs
RayTracingDriver::set_rays + run (include/RayTracingDriver.hpp:119-150) for a
whole ray batch on the GPU and returns the per-ray records.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L


@dataclass
class TraceResult:
    status: np.ndarray       # int8 [n]: 0 escaped, 1 trapped, 2 error
    err_kind: np.ndarray     # int8 [n]
    bounces: np.ndarray      # int32 [n]
    err_patch: np.ndarray    # int32 [n]
    hit_patch: np.ndarray    # int32 [n, K]
    hit_xyz: np.ndarray      # float64 [n, K, 3]
    iters: np.ndarray        # uint8 [n, K]
    final_od: np.ndarray     # float64 [n, 6]
    counters: np.ndarray     # int64 [16]
    hit_count: Optional[np.ndarray] = None
    ray_work: Optional[np.ndarray] = None   # uint32 [n, 4]

    def counter(self, name: str) -> int:
        return int(self.counters[L.CTR[name]])


def prepare_patches(sc, coarse: bool = False):
    """Host-only patch preparation (bbrt_prepare_patches): box, centroid, max sagitta [, coarse grid, cells]."""
    n = sc.n_patches
    box = np.zeros((n, 6)); cent = np.zeros((n, 3)); ms = np.zeros(n)
    cx = np.zeros((n, 100 * 100, 3)) if coarse else None
    cell = np.zeros((n, 99 * 99, 4)) if coarse else None
    rc = L.lib().bbrt_prepare_patches(C.byref(L.patch_set(sc)), L.ptr(box, L._d), L.ptr(cent, L._d), L.ptr(ms, L._d),
                                      L.ptr(cx, L._d) if coarse else None, L.ptr(cell, L._d) if coarse else None)
    L.check(rc, "bbrt_prepare_patches")
    return (box, cent, ms, cx, cell) if coarse else (box, cent, ms)


def build_bvh(box, cent, leaf_threshold: int = 4):
    box = np.ascontiguousarray(box, dtype=np.float64); cent = np.ascontiguousarray(cent, dtype=np.float64)
    n = cent.shape[0]
    nb = np.zeros((2 * n, 6)); r = np.zeros(2 * n, np.int32); s = np.zeros(2 * n, np.int32)
    c = np.zeros(2 * n, np.int32); a = np.zeros(2 * n, np.uint8); lo = np.zeros(n, np.int32)
    nn = L.lib().bbrt_build_bvh(L.ptr(box, L._d), L.ptr(cent, L._d), C.c_int64(n), C.c_int32(leaf_threshold),
                                L.ptr(nb, L._d), L.ptr(r, L._i32), L.ptr(s, L._i32), L.ptr(c, L._i32),
                                L.ptr(a, L._u8), L.ptr(lo, L._i32))
    if nn < 0:
        raise L.BBRTError("bbrt_build_bvh: " + L.last_error())
    return dict(box=nb[:nn], right=r[:nn], start=s[:nn], count=c[:nn], axis=a[:nn], leaf=lo)


class GpuScene:
    """A scene resident on the GPU (bbrt_scene)."""

    def __init__(self, sc):
        self.n_patches = sc.n_patches
        h = C.c_void_p()
        L.check(L.lib().bbrt_scene_create(C.byref(L.patch_set(sc)), C.byref(h)), "bbrt_scene_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().bbrt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        diag = C.c_double(); ms = C.c_double(); nn = C.c_int64(); np_ = C.c_int64(); by = C.c_int64()
        L.check(L.lib().bbrt_scene_info(self.h, C.byref(diag), C.byref(ms), C.byref(nn), C.byref(np_), C.byref(by)),
                "bbrt_scene_info")
        return dict(diag=diag.value, max_patch_sagitta=ms.value, n_nodes=nn.value, n_patches=np_.value,
                    device_bytes=by.value)

    def patch_info(self):
        n = self.n_patches
        box = np.zeros((n, 6)); cent = np.zeros((n, 3)); ms = np.zeros(n)
        L.check(L.lib().bbrt_scene_patch_info(self.h, L.ptr(box, L._d), L.ptr(cent, L._d), L.ptr(ms, L._d)),
                "bbrt_scene_patch_info")
        return box, cent, ms

    def bvh(self):
        m = self.info()["n_nodes"]
        nb = np.zeros((m, 6)); r = np.zeros(m, np.int32); s = np.zeros(m, np.int32)
        c = np.zeros(m, np.int32); a = np.zeros(m, np.uint8); lo = np.zeros(self.n_patches, np.int32)
        L.check(L.lib().bbrt_scene_bvh(self.h, L.ptr(nb, L._d), L.ptr(r, L._i32), L.ptr(s, L._i32), L.ptr(c, L._i32),
                                       L.ptr(a, L._u8), L.ptr(lo, L._i32)), "bbrt_scene_bvh")
        return dict(box=nb, right=r, start=s, count=c, axis=a, leaf=lo)

    def coarse(self, k: int):
        xyz = np.zeros((100 * 100, 3)); sag = np.zeros(99 * 99); cn = np.zeros((99 * 99, 3))
        L.check(L.lib().bbrt_scene_coarse(self.h, C.c_int64(k), L.ptr(xyz, L._d), L.ptr(sag, L._d), L.ptr(cn, L._d)),
                "bbrt_scene_coarse")
        return xyz, sag, cn

    def trace(self, origins, dirs, max_bounces: int, mode: int = L.NEWTON_CONVERGED,
              record_k: Optional[int] = None, flags: int = 0, hit_count: bool = False, block: int = 0,
              ray_work: bool = False, waves_per_eu: int = 0) -> TraceResult:
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        if o.shape != d.shape:
            raise ValueError("origins and dirs must have the same shape")
        n = o.shape[0]
        K = max(1, max_bounces if record_k is None else record_k)
        res = TraceResult(np.zeros(n, np.int8), np.zeros(n, np.int8), np.zeros(n, np.int32), np.zeros(n, np.int32),
                          np.zeros((n, K), np.int32), np.zeros((n, K, 3)), np.zeros((n, K), np.uint8),
                          np.zeros((n, 6)), np.zeros(L.NCOUNTERS, np.int64),
                          np.zeros(self.n_patches, np.int64) if hit_count else None,
                          np.zeros((n, 4), np.uint32) if ray_work else None)
        p = L.Params(mode, L.SEED_TRIANGLE, max_bounces, K, flags, block, waves_per_eu)
        out = L.Outputs(L.vptr(res.status), L.vptr(res.err_kind), L.vptr(res.bounces), L.vptr(res.err_patch),
                        L.vptr(res.hit_patch), L.vptr(res.hit_xyz), L.vptr(res.iters), L.vptr(res.final_od),
                        L.vptr(res.hit_count), L.vptr(res.ray_work))
        rc = L.lib().bbrt_trace(self.h, L.ptr(o, L._d), L.ptr(d, L._d), C.c_int64(n), C.byref(p), C.byref(out),
                                L.ptr(res.counters, L._i64))
        L.check(rc, "bbrt_trace")
        return res

    def trace_device(self, origins_ptr: int, dirs_ptr: int, n: int, max_bounces: int, out_ptrs: dict,
                     mode: int = L.NEWTON_CONVERGED, record_k: int = 0, flags: int = 0, stream: int = 0,
                     block: int = 0, waves_per_eu: int = 0):
        """Trace with device-resident inputs/outputs (raw device addresses). Returns (counters, kernel_ms[3])."""
        p = L.Params(mode, L.SEED_TRIANGLE, max_bounces, record_k, flags, block, waves_per_eu)
        out = L.Outputs(*[C.c_void_p(out_ptrs.get(k) or 0) if out_ptrs.get(k) else None
                          for k in ("status", "err_kind", "bounces", "err_patch", "hit_patch", "hit_xyz",
                                    "newton_iters", "final_od", "hit_count", "ray_work")])
        counters = np.zeros(L.NCOUNTERS, np.int64)
        kms = np.zeros(3)
        rc = L.lib().bbrt_trace_device(self.h, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr), C.c_int64(n),
                                       C.byref(p), C.byref(out), L.ptr(counters, L._i64),
                                       C.c_void_p(stream) if stream else None, L.ptr(kms, L._d))
        L.check(rc, "bbrt_trace_device")
        return counters, kms
