"""ctypes binding of libbbrt_hip.so (include/bbrt.h).

The library must be built in-tree (``python -m billardballraytracing3d_amd.build``
or ``__graft_entry__.build()``); loading fails loudly otherwise.  There is no
CPU fallback: every trace runs the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libbbrt_hip.so")

NEWTON_REF, NEWTON_CONVERGED, NEWTON_CONVERGED_FP32 = 0, 1, 2
SEED_TRIANGLE, SEED_LEARNED = 0, 1
ESCAPED, TRAPPED, ERROR = 0, 1, 2
ERR_NONE, ERR_SINGULAR, ERR_NOCONV, ERR_DOMAIN, ERR_BACKFACE, ERR_ZERO_DIR, ERR_STACK = range(7)
FLAG_NO_CULL, FLAG_COUNT_DIRECT_HIT, FLAG_TIME_KERNELS, FLAG_WAVEFRONT, FLAG_GRAPH = 1, 2, 4, 8, 16
FLAG_STAGE1_WAVE, FLAG_STAGE1_LANE = 32, 64
FLAG_GROUP_RAYS = 128
FLAG_SEED_FIRST = 256
SCENE_COMPACT, SCENE_DEVICE_BUILD = 1, 2
NCOUNTERS = 20
CTR = dict(rays=0, bounces=1, bvh_node_tests=2, candidates=3, cull_box_tests=4, mt_tests=5, newton_calls=6,
           newton_iters=7, newton_flops=8, polish_fail=9, launches=10, seed_learned=11, seed_fallback=12,
           mlp_evals=13, mlp_lanes=14, newton_iters_fp32=15, seed_miss=16)
MLP_NIN = 51
MLP_NPARAMS = 13520

# Every symbol include/bbrt.h declares.
EXPORTS = ("bbrt_scene_create", "bbrt_scene_destroy", "bbrt_scene_info", "bbrt_scene_patch_info", "bbrt_scene_bvh",
           "bbrt_scene_coarse", "bbrt_trace", "bbrt_trace_device", "bbrt_prepare_patches", "bbrt_build_bvh",
           "bbrt_last_error", "bbrt_mlp_create", "bbrt_mlp_destroy", "bbrt_mlp_get_params", "bbrt_mlp_set_params",
           "bbrt_mlp_forward", "bbrt_mlp_loss_grad", "bbrt_mlp_train", "bbrt_mlp_sample_pairs", "bbrt_mlp_features",
           "bbrt_scene_set_mlp", "bbrt_mlp_seed_ab", "bbrt_scene_create_batch", "bbrt_scene_batch_info",
           "bbrt_trace_batch", "bbrt_trace_batch_device", "bbrt_autotune", "bbrt_scene_create_ex",
           "bbrt_trace_device_async", "bbrt_trace_device_noalloc", "bbrt_moller_trumbore", "bbrt_triangles_create",
           "bbrt_triangles_destroy", "bbrt_triangles_intersect", "bbrt_scene_intersect", "bbrt_bounce_map",
           "bbrt_build_bvh_gpu")
NOALLOC_SCRATCH_BYTES = 64 << 20

_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_i8 = C.POINTER(C.c_int8)
_u8 = C.POINTER(C.c_uint8)
_f32 = C.POINTER(C.c_float)


class PatchSet(C.Structure):
    _fields_ = [("n_patches", C.c_int64), ("imax", _i32), ("jmax", _i32), ("u_off", _i64), ("v_off", _i64),
                ("f_off", _i64), ("unodes", _d), ("vnodes", _d), ("fields", _d)]


class Params(C.Structure):
    _fields_ = [("newton_mode", C.c_int32), ("seed_mode", C.c_int32), ("max_bounces", C.c_int32),
                ("record_k", C.c_int32), ("flags", C.c_uint32), ("block", C.c_int32), ("waves_per_eu", C.c_int32),
                ("grid_fraction", C.c_float), ("newton_batch", C.c_int32)]


class TuneEntry(C.Structure):
    _fields_ = [("block", C.c_int32), ("waves_per_eu", C.c_int32), ("newton_batch", C.c_int32),
                ("grid_fraction", C.c_float), ("ms", C.c_double)]


class Outputs(C.Structure):
    _fields_ = [("status", C.c_void_p), ("err_kind", C.c_void_p), ("bounces", C.c_void_p), ("err_patch", C.c_void_p),
                ("hit_patch", C.c_void_p), ("hit_xyz", C.c_void_p), ("newton_iters", C.c_void_p),
                ("final_od", C.c_void_p), ("hit_count", C.c_void_p), ("ray_work", C.c_void_p),
                ("seed_stats", C.c_void_p)]


class BounceMapParams(C.Structure):
    _fields_ = [("grid_n", C.c_int32), ("theta_n", C.c_int32), ("phi_n", C.c_int32), ("max_bounces", C.c_int32),
                ("newton_mode", C.c_int32), ("flags", C.c_uint32), ("cos_theta_max", C.c_double),
                ("cos_phi", C.POINTER(C.c_double)), ("sin_phi", C.POINTER(C.c_double)), ("eps", C.POINTER(C.c_double))]


class BounceMapOutputs(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("max_bounces", C.c_void_p), ("error_count", C.c_void_p),
                ("best_dir", C.c_void_p), ("best_od", C.c_void_p), ("launch_o", C.c_void_p), ("launch_d", C.c_void_p)]


class TrainConfig(C.Structure):
    _fields_ = [("n_samples", C.c_int64), ("seed", C.c_uint64), ("batch", C.c_int32), ("epochs", C.c_int32),
                ("lr", C.c_float), ("lr_final", C.c_float), ("val_frac", C.c_float), ("test_frac", C.c_float)]


class TrainReport(C.Structure):
    _fields_ = [("n_pairs", C.c_int64), ("n_train", C.c_int64), ("n_val", C.c_int64), ("n_test", C.c_int64),
                ("steps", C.c_int64), ("first_loss", C.c_float), ("final_loss", C.c_float), ("val_loss", C.c_float),
                ("test_loss", C.c_float), ("gen_ms", C.c_double), ("train_ms", C.c_double)]


class BBRTError(RuntimeError):
    pass


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise BBRTError(f"libbbrt_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.bbrt_last_error.restype = C.c_char_p
        L.bbrt_scene_create.argtypes = [C.POINTER(PatchSet), C.POINTER(C.c_void_p)]
        L.bbrt_scene_create.restype = C.c_int
        L.bbrt_scene_destroy.argtypes = [C.c_void_p]
        L.bbrt_scene_destroy.restype = None
        L.bbrt_scene_info.argtypes = [C.c_void_p, _d, _d, _i64, _i64, _i64]
        L.bbrt_scene_patch_info.argtypes = [C.c_void_p, _d, _d, _d]
        L.bbrt_scene_bvh.argtypes = [C.c_void_p, _d, _i32, _i32, _i32, _u8, _i32]
        L.bbrt_scene_coarse.argtypes = [C.c_void_p, C.c_int64, _d, _d, _d]
        L.bbrt_trace.argtypes = [C.c_void_p, _d, _d, C.c_int64, C.POINTER(Params), C.POINTER(Outputs), _i64]
        L.bbrt_trace_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(Params),
                                        C.POINTER(Outputs), _i64, C.c_void_p, _d]
        L.bbrt_prepare_patches.argtypes = [C.POINTER(PatchSet), _d, _d, _d, _d, _d]
        L.bbrt_build_bvh.argtypes = [_d, _d, C.c_int64, C.c_int32, _d, _i32, _i32, _i32, _u8, _i32]
        L.bbrt_build_bvh.restype = C.c_int64
        L.bbrt_build_bvh_gpu.argtypes = [_d, _d, C.c_int64, C.c_int32, _d, _i32, _i32, _i32, _u8, _i32]
        L.bbrt_build_bvh_gpu.restype = C.c_int64
        _f = C.POINTER(C.c_float)
        L.bbrt_mlp_create.argtypes = [C.c_uint64, C.POINTER(C.c_void_p)]
        L.bbrt_mlp_destroy.argtypes = [C.c_void_p]
        L.bbrt_mlp_destroy.restype = None
        L.bbrt_mlp_get_params.argtypes = [C.c_void_p, _f]
        L.bbrt_mlp_set_params.argtypes = [C.c_void_p, _f]
        L.bbrt_mlp_forward.argtypes = [C.c_void_p, _f, C.c_int64, _f]
        L.bbrt_mlp_loss_grad.argtypes = [C.c_void_p, _f, _f, C.c_int64, _f, _f]
        L.bbrt_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(TrainConfig), _f, C.POINTER(TrainReport)]
        L.bbrt_mlp_sample_pairs.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, _i32, _d, _d, _f, _f, _i8]
        L.bbrt_mlp_features.argtypes = [C.c_void_p, _i32, _d, _d, C.c_int64, _f, _d, _d]
        L.bbrt_scene_set_mlp.argtypes = [C.c_void_p, C.c_void_p]
        L.bbrt_mlp_seed_ab.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_int64, _i8, _u8, _u8, _f, _d]
        L.bbrt_scene_create_batch.argtypes = [C.POINTER(PatchSet), C.c_int32, C.POINTER(C.c_void_p)]
        L.bbrt_scene_batch_info.argtypes = [C.c_void_p, _i32, _i64, _i32, _d]
        L.bbrt_scene_create_ex.argtypes = [C.POINTER(PatchSet), C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]
        L.bbrt_trace_batch.argtypes = [C.c_void_p, _d, _d, _i64, C.POINTER(Params), C.POINTER(Outputs), _i64]
        L.bbrt_trace_batch_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _i64, C.POINTER(Params),
                                              C.POINTER(Outputs), _i64, C.c_void_p, _d]
        L.bbrt_trace_device_async.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, _i64, C.POINTER(Params),
                                              C.POINTER(Outputs), C.c_void_p, C.c_void_p]
        L.bbrt_trace_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(Params),
                                                C.POINTER(Outputs), C.c_void_p, C.c_void_p]
        L.bbrt_moller_trumbore.argtypes = [_d, _d, _d, C.c_int64, C.c_double, _i8, _d, _d, _d]
        L.bbrt_triangles_create.argtypes = [_d, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]
        L.bbrt_triangles_destroy.argtypes = [C.c_void_p]
        L.bbrt_triangles_destroy.restype = None
        L.bbrt_triangles_intersect.argtypes = [C.c_void_p, _d, _d, C.c_int64, C.c_int32, _i32, _d, _d, _d]
        L.bbrt_scene_intersect.argtypes = [C.c_void_p, _d, _d, C.c_int64, C.c_int32, C.c_int32, _i32, _d, _d, _d, _i8,
                                           _i32]
        L.bbrt_bounce_map.argtypes = [C.c_void_p, _i64, C.c_int64, C.POINTER(BounceMapParams),
                                      C.POINTER(BounceMapOutputs), _i64]
        L.bbrt_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _i64, C.c_int64, C.POINTER(Params),
                                    C.POINTER(Outputs), C.POINTER(Params), C.POINTER(TuneEntry), C.c_int32, _i32,
                                    C.c_void_p]
        _LIB = L
    return _LIB


def last_error() -> str:
    return lib().bbrt_last_error().decode()


def check(rc: int, what: str):
    if rc != 0:
        raise BBRTError(f"{what}: {last_error()}")


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def vptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def patch_set(sc) -> PatchSet:
    for name in ("imax", "jmax"):
        assert getattr(sc, name).dtype == np.int32 and getattr(sc, name).flags.c_contiguous
    for name in ("u_off", "v_off", "f_off"):
        assert getattr(sc, name).dtype == np.int64 and getattr(sc, name).flags.c_contiguous
    for name in ("unodes", "vnodes", "fields"):
        assert getattr(sc, name).dtype == np.float64 and getattr(sc, name).flags.c_contiguous
    ps = PatchSet()
    ps.n_patches = sc.n_patches
    ps.imax = ptr(sc.imax, _i32); ps.jmax = ptr(sc.jmax, _i32)
    ps.u_off = ptr(sc.u_off, _i64); ps.v_off = ptr(sc.v_off, _i64); ps.f_off = ptr(sc.f_off, _i64)
    ps.unodes = ptr(sc.unodes, _d); ps.vnodes = ptr(sc.vnodes, _d); ps.fields = ptr(sc.fields, _d)
    return ps
