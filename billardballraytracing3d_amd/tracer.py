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
    counters: np.ndarray     # int64 [NCOUNTERS]
    hit_count: Optional[np.ndarray] = None
    ray_work: Optional[np.ndarray] = None   # uint32 [n, 4]
    seed_stats: Optional[np.ndarray] = None # uint32 [n, 2]: learned roots kept / fallbacks

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


def build_bvh(box, cent, leaf_threshold: int = 4, gpu: bool = False):
    """BVH(objects, leaf_threshold) over boxes [n][6] and centroids [n][3]: built on the host
    (bbrt_build_bvh) or on the GPU (bbrt_build_bvh_gpu, the same tree)."""
    box = np.ascontiguousarray(box, dtype=np.float64); cent = np.ascontiguousarray(cent, dtype=np.float64)
    n = cent.shape[0]
    nb = np.zeros((2 * n, 6)); r = np.zeros(2 * n, np.int32); s = np.zeros(2 * n, np.int32)
    c = np.zeros(2 * n, np.int32); a = np.zeros(2 * n, np.uint8); lo = np.zeros(n, np.int32)
    fn = L.lib().bbrt_build_bvh_gpu if gpu else L.lib().bbrt_build_bvh
    nn = fn(L.ptr(box, L._d), L.ptr(cent, L._d), C.c_int64(n), C.c_int32(leaf_threshold), L.ptr(nb, L._d),
            L.ptr(r, L._i32), L.ptr(s, L._i32), L.ptr(c, L._i32), L.ptr(a, L._u8), L.ptr(lo, L._i32))
    if nn < 0:
        raise L.BBRTError(("bbrt_build_bvh_gpu: " if gpu else "bbrt_build_bvh: ") + L.last_error())
    return dict(box=nb[:nn], right=r[:nn], start=s[:nn], count=c[:nn], axis=a[:nn], leaf=lo)


def moller_trumbore(tri, origins, dirs, eps: float = 1e-9):
    """moller_trumbore (MollerTrumbore.hpp:32-64) on (ray, triangle) pairs on the GPU: tri [n][9] (v0, v1, v2),
    origins / dirs [n][3] -> (hit [n] bool, t [n], point [n][3], normal [n][3])."""
    tri = np.ascontiguousarray(tri, dtype=np.float64).reshape(-1, 9)
    o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    n = tri.shape[0]
    if o.shape[0] != n or d.shape[0] != n:
        raise ValueError("tri, origins and dirs must have the same length")
    hit = np.zeros(n, np.int8); t = np.zeros(n); p = np.zeros((n, 3)); nn = np.zeros((n, 3))
    L.check(L.lib().bbrt_moller_trumbore(L.ptr(tri, L._d), L.ptr(o, L._d), L.ptr(d, L._d), C.c_int64(n), C.c_double(eps),
                                         L.ptr(hit, L._i8), L.ptr(t, L._d), L.ptr(p, L._d), L.ptr(nn, L._d)),
            "bbrt_moller_trumbore")
    return hit.astype(bool), t, p, nn


class Triangles:
    """A set of flat triangles with its BVH on the GPU (bbrt_triangles): BVH::intersect or brute_force_intersect
    (BVH.hpp:75-128, BruteForce.hpp:7-15) for many rays at once."""

    def __init__(self, tri, leaf_threshold: int = 4):
        tri = np.ascontiguousarray(tri, dtype=np.float64).reshape(-1, 9)
        h = C.c_void_p()
        L.check(L.lib().bbrt_triangles_create(L.ptr(tri, L._d), C.c_int64(tri.shape[0]), C.c_int32(leaf_threshold),
                                              C.byref(h)), "bbrt_triangles_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().bbrt_triangles_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def intersect(self, origins, dirs, brute_force: bool = False):
        """-> (object [n] (-1 miss), t [n], point [n][3], normal [n][3])."""
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        obj = np.zeros(n, np.int32); t = np.zeros(n); p = np.zeros((n, 3)); nn = np.zeros((n, 3))
        L.check(L.lib().bbrt_triangles_intersect(self.h, L.ptr(o, L._d), L.ptr(d, L._d), C.c_int64(n),
                                                 C.c_int32(int(brute_force)), L.ptr(obj, L._i32), L.ptr(t, L._d),
                                                 L.ptr(p, L._d), L.ptr(nn, L._d)), "bbrt_triangles_intersect")
        return obj, t, p, nn


def _params(mode, seed, max_bounces, record_k, flags, block, waves_per_eu, newton_batch, grid_fraction=0.0):
    return L.Params(newton_mode=mode, seed_mode=seed, max_bounces=max_bounces, record_k=record_k, flags=flags,
                    block=block, waves_per_eu=waves_per_eu, grid_fraction=grid_fraction, newton_batch=newton_batch)


def _outputs(out_ptrs: dict):
    return L.Outputs(*[C.c_void_p(out_ptrs.get(k) or 0) if out_ptrs.get(k) else None
                       for k in ("status", "err_kind", "bounces", "err_patch", "hit_patch", "hit_xyz",
                                 "newton_iters", "final_od", "hit_count", "ray_work", "seed_stats")])


class GpuScene:
    """A scene resident on the GPU (bbrt_scene); GpuScene.batch(scenes) holds several sub-scenes."""

    def __init__(self, sc=None):
        self.n_scenes = 1
        if sc is None:
            return
        self.n_patches = sc.n_patches
        h = C.c_void_p()
        L.check(L.lib().bbrt_scene_create(C.byref(L.patch_set(sc)), C.byref(h)), "bbrt_scene_create")
        self.h = h
        self.patch_begin = np.array([0, sc.n_patches], np.int64)

    @classmethod
    def batch(cls, scenes, compact: bool = False, device_build: bool = False):
        """Several independent scenes in one handle (bbrt_scene_create_ex; compact: coarse cells on the fly;
        device_build: the compact scene's per-cell build work on the GPU)."""
        self = cls(None)
        sets = (L.PatchSet * len(scenes))(*[L.patch_set(sc) for sc in scenes])
        self._keep = scenes
        h = C.c_void_p()
        fl = (L.SCENE_COMPACT if compact else 0) | (L.SCENE_DEVICE_BUILD if device_build else 0)
        L.check(L.lib().bbrt_scene_create_ex(sets, C.c_int32(len(scenes)), C.c_uint32(fl), C.byref(h)),
                "bbrt_scene_create_ex")
        self.h = h
        self.n_scenes = len(scenes)
        self.patch_begin = np.zeros(self.n_scenes + 1, np.int64)
        ns = C.c_int32()
        L.check(L.lib().bbrt_scene_batch_info(self.h, C.byref(ns), L.ptr(self.patch_begin, L._i64), None, None),
                "bbrt_scene_batch_info")
        self.n_patches = int(self.patch_begin[-1])
        return self

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

    def bounce_map(self, patches, grid_n: int, theta_n: int, phi_n: int, cone_angle_deg: float = 180.0,
                   max_bounces: int = 50, mode: int = L.NEWTON_REF, launch_rays: bool = False):
        """BounceMapDriver::run (BounceMapDriver.hpp:103-259) on the device for the given patch indices
        (bbrt_bounce_map).  The launch nudge of each patch is the driver's self_intersect_epsilon,
        max(1e-9 * diagonal, 5 * its max sagitta); cos(theta_max) and the phi table use the C library's cos/sin.
        Returns a dict of grid [m][n*n][9], max_bounces, error_count, best_dir [m][n*n], best_od [m][n*n][6],
        counters, and with launch_rays the launch origins / directions."""
        import math
        pid = np.ascontiguousarray(patches, dtype=np.int64)
        m = pid.size
        info = self.info()
        _, _, ms = self.patch_info()
        # (indices out of range are left to the library to refuse)
        sag = np.array([ms[k] if 0 <= k < ms.size else 0.0 for k in pid])
        eps = np.array([s_ * 5.0 if not (5.0 * s_ < 1e-9 * info["diag"]) else 1e-9 * info["diag"] for s_ in sag])
        kTwoPi = 6.283185307179586476925286766559
        cphi = np.array([math.cos((j + 0.5) * kTwoPi / phi_n) for j in range(phi_n)])
        sphi = np.array([math.sin((j + 0.5) * kTwoPi / phi_n) for j in range(phi_n)])
        bp = L.BounceMapParams(grid_n, theta_n, phi_n, max_bounces, mode, 0,
                               math.cos(cone_angle_deg / 2.0 * kTwoPi / 360.0), L.ptr(cphi, L._d), L.ptr(sphi, L._d),
                               L.ptr(eps, L._d))
        npts = m * grid_n * grid_n
        nr = npts * theta_n * phi_n
        res = dict(grid=np.zeros((m, grid_n * grid_n, 9)), max_bounces=np.zeros((m, grid_n * grid_n), np.int32),
                   error_count=np.zeros((m, grid_n * grid_n), np.int32), best_dir=np.zeros((m, grid_n * grid_n), np.int32),
                   best_od=np.zeros((m, grid_n * grid_n, 6)), counters=np.zeros(L.NCOUNTERS, np.int64), eps=eps)
        if launch_rays:
            res["launch_o"] = np.zeros((nr, 3)); res["launch_d"] = np.zeros((nr, 3))
        out = L.BounceMapOutputs(*[L.vptr(res.get(k)) for k in ("grid", "max_bounces", "error_count", "best_dir",
                                                                 "best_od", "launch_o", "launch_d")])
        L.check(L.lib().bbrt_bounce_map(self.h, L.ptr(pid, L._i64), C.c_int64(m), C.byref(bp), C.byref(out),
                                        L.ptr(res["counters"], L._i64)), "bbrt_bounce_map")
        return res

    def intersect(self, origins, dirs, mode: int = L.NEWTON_REF, brute_force: bool = False):
        """BVH::intersect (or brute_force_intersect) over the patches for many rays, directions as given
        (bbrt_scene_intersect) -> dict of patch, t, point, normal, err_kind, err_patch."""
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        patch = np.zeros(n, np.int32); t = np.zeros(n); p = np.zeros((n, 3)); nn = np.zeros((n, 3))
        ek = np.zeros(n, np.int8); ep = np.zeros(n, np.int32)
        L.check(L.lib().bbrt_scene_intersect(self.h, L.ptr(o, L._d), L.ptr(d, L._d), C.c_int64(n), C.c_int32(mode),
                                             C.c_int32(int(brute_force)), L.ptr(patch, L._i32), L.ptr(t, L._d),
                                             L.ptr(p, L._d), L.ptr(nn, L._d), L.ptr(ek, L._i8), L.ptr(ep, L._i32)),
                "bbrt_scene_intersect")
        return dict(patch=patch, t=t, point=p, normal=nn, err_kind=ek, err_patch=ep)

    def trace(self, origins, dirs, max_bounces: int, mode: int = L.NEWTON_CONVERGED,
              record_k: Optional[int] = None, flags: int = 0, hit_count: bool = False, block: int = 0,
              ray_work: bool = False, waves_per_eu: int = 0, seed: int = L.SEED_TRIANGLE,
              newton_batch: int = 0, grid_fraction: float = 0.0, ray_begin=None,
              seed_stats: bool = False) -> TraceResult:
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
                          np.zeros((n, 4), np.uint32) if ray_work else None,
                          np.zeros((n, 2), np.uint32) if seed_stats else None)
        p = _params(mode, seed, max_bounces, K, flags, block, waves_per_eu, newton_batch, grid_fraction)
        out = L.Outputs(L.vptr(res.status), L.vptr(res.err_kind), L.vptr(res.bounces), L.vptr(res.err_patch),
                        L.vptr(res.hit_patch), L.vptr(res.hit_xyz), L.vptr(res.iters), L.vptr(res.final_od),
                        L.vptr(res.hit_count), L.vptr(res.ray_work), L.vptr(res.seed_stats))
        if ray_begin is None:
            rc = L.lib().bbrt_trace(self.h, L.ptr(o, L._d), L.ptr(d, L._d), C.c_int64(n), C.byref(p), C.byref(out),
                                    L.ptr(res.counters, L._i64))
            L.check(rc, "bbrt_trace")
        else:
            rb = np.ascontiguousarray(ray_begin, dtype=np.int64)
            if rb.size != self.n_scenes + 1 or rb[-1] != n:
                raise ValueError("ray_begin must have n_scenes + 1 entries ending at the ray count")
            rc = L.lib().bbrt_trace_batch(self.h, L.ptr(o, L._d), L.ptr(d, L._d), L.ptr(rb, L._i64), C.byref(p),
                                          C.byref(out), L.ptr(res.counters, L._i64))
            L.check(rc, "bbrt_trace_batch")
        return res

    def trace_device(self, origins_ptr: int, dirs_ptr: int, n: int, max_bounces: int, out_ptrs: dict,
                     mode: int = L.NEWTON_CONVERGED, record_k: int = 0, flags: int = 0, stream: int = 0,
                     block: int = 0, waves_per_eu: int = 0, seed: int = L.SEED_TRIANGLE, newton_batch: int = 0,
                     grid_fraction: float = 0.0, ray_begin=None):
        """Trace with device-resident inputs/outputs (raw device addresses). Returns (counters, kernel_ms[3]).
        With ray_begin (host array, n_scenes + 1 entries) the rays are traced per sub-scene (bbrt_trace_batch_device)."""
        p = _params(mode, seed, max_bounces, record_k, flags, block, waves_per_eu, newton_batch, grid_fraction)
        out = _outputs(out_ptrs)
        counters = np.zeros(L.NCOUNTERS, np.int64)
        kms = np.zeros(3)
        st = C.c_void_p(stream) if stream else None
        if ray_begin is None:
            rc = L.lib().bbrt_trace_device(self.h, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr), C.c_int64(n),
                                           C.byref(p), C.byref(out), L.ptr(counters, L._i64), st, L.ptr(kms, L._d))
            L.check(rc, "bbrt_trace_device")
        else:
            rb = np.ascontiguousarray(ray_begin, dtype=np.int64)
            rc = L.lib().bbrt_trace_batch_device(self.h, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr),
                                                 L.ptr(rb, L._i64), C.byref(p), C.byref(out),
                                                 L.ptr(counters, L._i64), st, L.ptr(kms, L._d))
            L.check(rc, "bbrt_trace_batch_device")
        return counters, kms

    def trace_device_async(self, origins_ptr: int, dirs_ptr: int, n: int, max_bounces: int, out_ptrs: dict,
                           counters_ptr: int = 0, mode: int = L.NEWTON_CONVERGED, record_k: int = 0, flags: int = 0,
                           stream: int = 0, block: int = 0, waves_per_eu: int = 0, seed: int = L.SEED_TRIANGLE,
                           newton_batch: int = 0, grid_fraction: float = 0.0, ray_begin=None):
        """bbrt_trace_device_async: enqueue the trace on `stream` and return at once.  counters_ptr: 0 or a
        device address of NCOUNTERS int64, written on the stream.  Traces on distinct streams may overlap."""
        p = _params(mode, seed, max_bounces, record_k, flags, block, waves_per_eu, newton_batch, grid_fraction)
        out = _outputs(out_ptrs)
        rb = None if ray_begin is None else np.ascontiguousarray(ray_begin, dtype=np.int64)
        rc = L.lib().bbrt_trace_device_async(self.h, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr), C.c_int64(n),
                                             L.ptr(rb, L._i64) if rb is not None else None, C.byref(p),
                                             C.byref(out), C.c_void_p(counters_ptr) if counters_ptr else None,
                                             C.c_void_p(stream) if stream else None)
        L.check(rc, "bbrt_trace_device_async")

    def trace_device_noalloc(self, origins_ptr: int, dirs_ptr: int, n: int, max_bounces: int, out_ptrs: dict,
                             counters_ptr: int = 0, mode: int = L.NEWTON_CONVERGED, record_k: int = 0, flags: int = 0,
                             stream: int = 0, block: int = 0, waves_per_eu: int = 0, seed: int = L.SEED_TRIANGLE,
                             newton_batch: int = 0, grid_fraction: float = 0.0):
        """bbrt_trace_device_noalloc: bbrt_trace_device_async with the trace's scratch in the library's fixed
        per-GPU array (L.NOALLOC_SCRATCH_BYTES) instead of device memory allocated for it.  Enqueues on `stream`
        and returns; calls that overlap on one GPU share that array and race."""
        p = _params(mode, seed, max_bounces, record_k, flags, block, waves_per_eu, newton_batch, grid_fraction)
        out = _outputs(out_ptrs)
        rc = L.lib().bbrt_trace_device_noalloc(self.h, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr), C.c_int64(n),
                                               C.byref(p), C.byref(out),
                                               C.c_void_p(counters_ptr) if counters_ptr else None,
                                               C.c_void_p(stream) if stream else None)
        L.check(rc, "bbrt_trace_device_noalloc")

    def autotune(self, origins_ptr: int, dirs_ptr: int, n: int, max_bounces: int, out_ptrs: dict,
                 mode: int = L.NEWTON_CONVERGED, record_k: int = 0, flags: int = 0, stream: int = 0,
                 seed: int = L.SEED_TRIANGLE, ray_begin=None):
        """bbrt_autotune: returns (best launch shape as a dict, list of timed candidates)."""
        base = _params(mode, seed, max_bounces, record_k, flags, 0, 0, 0)
        best = L.Params()
        table = (L.TuneEntry * 64)()
        cnt = C.c_int32()
        rb = None if ray_begin is None else np.ascontiguousarray(ray_begin, dtype=np.int64)
        rc = L.lib().bbrt_autotune(self.h, C.c_void_p(origins_ptr), C.c_void_p(dirs_ptr),
                                   L.ptr(rb, L._i64) if rb is not None else None, C.c_int64(n), C.byref(base),
                                   C.byref(_outputs(out_ptrs)), C.byref(best), table, C.c_int32(64), C.byref(cnt),
                                   C.c_void_p(stream) if stream else None)
        L.check(rc, "bbrt_autotune")
        pick = dict(block=best.block, waves_per_eu=best.waves_per_eu, newton_batch=best.newton_batch,
                    grid_fraction=best.grid_fraction)
        rows = [dict(block=t.block, waves_per_eu=t.waves_per_eu, newton_batch=t.newton_batch,
                     grid_fraction=round(t.grid_fraction, 3), ms=t.ms) for t in table[:cnt.value]]
        return pick, rows

    # ------------------------------------------------------------------
    # learned seed
    # ------------------------------------------------------------------
    def set_mlp(self, mlp: "Mlp"):
        """Attach a snapshot of the network (bbrt_scene_set_mlp); SEED_LEARNED traces use it."""
        L.check(L.lib().bbrt_scene_set_mlp(self.h, mlp.h), "bbrt_scene_set_mlp")

    def sample_pairs(self, seed: int, first: int, n: int):
        """(patch, origin, dir, x[51], target[3], valid) of training pairs first .. first+n-1."""
        patch = np.zeros(n, np.int32); o = np.zeros((n, 3)); d = np.zeros((n, 3))
        x = np.zeros((n, L.MLP_NIN), np.float32); t = np.zeros((n, 3), np.float32); v = np.zeros(n, np.int8)
        L.check(L.lib().bbrt_mlp_sample_pairs(self.h, C.c_uint64(seed), C.c_int64(first), C.c_int64(n),
                                              L.ptr(patch, L._i32), L.ptr(o, L._d), L.ptr(d, L._d),
                                              L.ptr(x, L._f32), L.ptr(t, L._f32), L.ptr(v, L._i8)),
                "bbrt_mlp_sample_pairs")
        return patch, o, d, x, t, v.astype(bool)

    def mlp_features(self, patch, origins, dirs, frame: bool = False):
        patch = np.ascontiguousarray(patch, dtype=np.int32)
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = patch.size
        x = np.zeros((n, L.MLP_NIN), np.float32); tc = np.zeros((n, 3))
        fr = np.zeros((self.n_patches, 24)) if frame else None
        L.check(L.lib().bbrt_mlp_features(self.h, L.ptr(patch, L._i32), L.ptr(o, L._d), L.ptr(d, L._d),
                                          C.c_int64(n), L.ptr(x, L._f32), L.ptr(tc, L._d),
                                          L.ptr(fr, L._d) if frame else None), "bbrt_mlp_features")
        return (x, tc, fr) if frame else (x, tc)

    def seed_ab(self, seed: int, first: int, n: int):
        """Triangle seed vs learned seed on sampled pairs (bbrt_mlp_seed_ab)."""
        st = np.zeros(2 * n, np.int8); it = np.zeros(2 * n, np.uint8); ac = np.zeros(n, np.uint8)
        se = np.zeros(2 * n, np.float32); pd = np.zeros(n)
        L.check(L.lib().bbrt_mlp_seed_ab(self.h, C.c_uint64(seed), C.c_int64(first), C.c_int64(n),
                                         L.ptr(st, L._i8), L.ptr(it, L._u8), L.ptr(ac, L._u8), L.ptr(se, L._f32),
                                         L.ptr(pd, L._d)), "bbrt_mlp_seed_ab")
        return dict(status_tri=st[:n], status_learned=st[n:], iters_tri=it[:n], iters_learned=it[n:],
                    accepted=ac.astype(bool), seed_err_tri=se[:n], seed_err_learned=se[n:], point_diff=pd)


class Mlp:
    """The learned-seed network (bbrt_mlp): GPU forward, gradients and training."""

    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        L.check(L.lib().bbrt_mlp_create(C.c_uint64(seed), C.byref(h)), "bbrt_mlp_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().bbrt_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(L.MLP_NPARAMS, np.float32)
        L.check(L.lib().bbrt_mlp_get_params(self.h, L.ptr(p, L._f32)), "bbrt_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != L.MLP_NPARAMS:
            raise ValueError("wrong parameter count")
        L.check(L.lib().bbrt_mlp_set_params(self.h, L.ptr(p, L._f32)), "bbrt_mlp_set_params")

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.MLP_NIN)
        y = np.zeros((x.shape[0], 3), np.float32)
        L.check(L.lib().bbrt_mlp_forward(self.h, L.ptr(x, L._f32), C.c_int64(x.shape[0]), L.ptr(y, L._f32)),
                "bbrt_mlp_forward")
        return y

    def loss_grad(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, L.MLP_NIN)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, 3)
        g = np.zeros(L.MLP_NPARAMS, np.float32); loss = C.c_float()
        L.check(L.lib().bbrt_mlp_loss_grad(self.h, L.ptr(x, L._f32), L.ptr(t, L._f32), C.c_int64(x.shape[0]),
                                           C.byref(loss), L.ptr(g, L._f32)), "bbrt_mlp_loss_grad")
        return loss.value, g

    def train(self, scene: GpuScene, n_samples: int, seed: int = 1, batch: int = 16384, epochs: int = 4,
              lr: float = 2e-3, lr_final: float = 2e-5, val_frac: float = 0.1, test_frac: float = 0.1):
        cfg = L.TrainConfig(n_samples, seed, batch, epochs, lr, lr_final, val_frac, test_frac)
        rep = L.TrainReport()
        n_train_max = int(n_samples * (1.0 - val_frac - test_frac)) + 1
        curve = np.zeros(max(1, (n_train_max // batch) * epochs + 1), np.float32)
        L.check(L.lib().bbrt_mlp_train(self.h, scene.h, C.byref(cfg), L.ptr(curve, L._f32), C.byref(rep)),
                "bbrt_mlp_train")
        report = {k: getattr(rep, k) for k, _ in L.TrainReport._fields_}
        return report, curve[:rep.steps]
