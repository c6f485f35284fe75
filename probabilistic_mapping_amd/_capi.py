"""ctypes binding of ``libpm_hip.so`` (declared in ``include/pm_capi.h``).

The library is built in-tree by ``probabilistic_mapping_amd.build``; loading fails loudly if it
is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpm_hip.so")

PM_OK = 0
PM_ERR_INVALID_ARGUMENT = 1
PM_ERR_OUT_OF_MEMORY = 2
PM_ERR_LAUNCH = 3
PM_ERR_INVALID_STATE = 4
PM_ERR_DEVICE = 5

PM_COST_F32 = 0
PM_COST_F16 = 1
PM_COST_BF16 = 2

PM_SGM_EXACT = 0
PM_SGM_COMPACT = 1
PM_SGM_SWEEP = 2

PM_MAX_PLANES = 256
PM_MAX_SOURCES = 16

STATUS_NAMES = {0: "PM_OK", 1: "PM_ERR_INVALID_ARGUMENT", 2: "PM_ERR_OUT_OF_MEMORY",
                3: "PM_ERR_LAUNCH", 4: "PM_ERR_INVALID_STATE", 5: "PM_ERR_DEVICE"}


class PmError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Shape(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("planes", ctypes.c_int32)]


class Homography(ctypes.Structure):
    _fields_ = [("h", ctypes.c_float * 9), ("t", ctypes.c_float * 3)]


class SgmParams(ctypes.Structure):
    _fields_ = [("pi1", ctypes.c_float), ("pi2", ctypes.c_float), ("tau_so", ctypes.c_float),
                ("sgm_q1", ctypes.c_float), ("sgm_q2", ctypes.c_float), ("paths", ctypes.c_int32)]


class Camera(ctypes.Structure):
    _fields_ = [("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float),
                ("cy", ctypes.c_float)]


class FusionState(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_void_p), ("beta", ctypes.c_void_p), ("mu", ctypes.c_void_p),
                ("sig", ctypes.c_void_p)]


PM_MAX_BATCH = 64


class Keyframe(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("planes", ctypes.c_int32), ("num_sources", ctypes.c_int32),
                ("idep_min", ctypes.c_float), ("dep_sample", ctypes.c_float),
                ("ref", ctypes.c_void_p), ("src_quads", ctypes.c_void_p * 16),
                ("homs", Homography * 16), ("cost", ctypes.c_void_p), ("agg", ctypes.c_void_p),
                ("depth", ctypes.c_void_p)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("cost_tile_rows", ctypes.c_int32), ("cost_stage_planes", ctypes.c_int32),
                ("sgm_prefetch", ctypes.c_int32), ("cost_blocks_per_cu", ctypes.c_int32),
                ("sgm_blocks_per_cu", ctypes.c_int32), ("cost_plane_split", ctypes.c_int32),
                ("sgm_rows", ctypes.c_int32)]


class Knobs(ctypes.Structure):
    _fields_ = [("pi1", ctypes.c_float), ("pi2", ctypes.c_float), ("tau_so", ctypes.c_float),
                ("sgm_q1", ctypes.c_float), ("sgm_q2", ctypes.c_float),
                ("var_scale", ctypes.c_float), ("case_mode", ctypes.c_int32),
                ("idep_min", ctypes.c_float), ("sgm_paths", ctypes.c_int32),
                ("median", ctypes.c_int32)]


class DiffParams(ctypes.Structure):
    _fields_ = [("temperature", ctypes.c_float), ("cost_scale", ctypes.c_float),
                ("recompute", ctypes.c_int32)]


class DriverConfig(ctypes.Structure):
    _fields_ = [("agg_num", ctypes.c_int32), ("interval", ctypes.c_int32),
                ("planes", ctypes.c_int32), ("fx", ctypes.c_double), ("fy", ctypes.c_double),
                ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("dep_sample", ctypes.c_float),
                ("frame_ring", ctypes.c_int32)]


_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32 = ctypes.c_int32

# name -> (restype, argtypes); every symbol declared in include/pm_capi.h
SIGNATURES = {
    "pm_last_error": (ctypes.c_char_p, []),
    "pm_version": (ctypes.c_char_p, []),
    "pm_homography_setup": (_i32, [_dp, _dp, _dp, _dp, _dp, ctypes.POINTER(Homography)]),
    "pm_make_quad": (_i32, [_vp, _vp, _i32, _i32, _vp]),
    "pm_calc_cost": (_i32, [Shape, _vp, _i32, ctypes.POINTER(_vp), ctypes.POINTER(Homography),
                            ctypes.c_float, ctypes.c_float, _vp, _i32, _vp]),
    "pm_sgm": (_i32, [Shape, _vp, _vp, _i32, SgmParams, _vp, _vp]),
    "pm_sgm_filter": (_i32, [Shape, _vp, _vp, _i32, SgmParams, ctypes.c_float, ctypes.c_float,
                             ctypes.c_float, _vp, _vp, _vp]),
    "pm_sgm_workspace_size": (_i32, [Shape, _i32, ctypes.POINTER(ctypes.c_uint64)]),
    "pm_sgm_filter_paths": (_i32, [Shape, _vp, _vp, _i32, SgmParams, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, _vp, ctypes.c_uint64, _vp,
                                   _vp, _vp]),
    "pm_sgm_excess_workspace_size": (_i32, [Shape, _i32, ctypes.POINTER(ctypes.c_uint64)]),
    "pm_sgm_filter_excess": (_i32, [Shape, _vp, _vp, _i32, SgmParams, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float, _vp, ctypes.c_uint64, _vp,
                                    _vp, _vp]),
    "pm_filter_cost": (_i32, [Shape, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp,
                              _vp]),
    "pm_wta": (_i32, [Shape, _vp, _i32, ctypes.c_float, ctypes.c_float, _vp, _vp]),
    "pm_median3x3": (_i32, [_vp, _vp, _i32, _i32, _vp]),
    "pm_depth_fuse": (_i32, [_i32, _i32, Camera, _fp, _fp, _vp, _vp, _vp, _vp, _vp,
                             FusionState, FusionState, _vp]),
    "pm_zbuf_reset": (_i32, [_vp, ctypes.c_size_t, _vp]),
    "pm_depth_fuse_warp": (_i32, [_i32, _i32, Camera, _fp, _fp, FusionState, _vp, _vp, _vp]),
    "pm_depth_fuse_step": (_i32, [_i32, _i32, Camera, _vp, _i32, _vp, _vp, _vp, _vp,
                                  FusionState, FusionState, _fp, _fp, _vp, _vp, _vp]),
    "pm_depth_fuse_noalloc": (_i32, [_i32, _i32, Camera, _fp, _fp, _vp, _vp, _vp, FusionState,
                                     FusionState, _vp]),
    "pm_fill_f32": (_i32, [_vp, ctypes.c_size_t, ctypes.c_float, _vp]),
    "pm_batch_create": (_i32, [ctypes.POINTER(Keyframe), _i32, _i32, SgmParams, ctypes.c_float,
                               _i32, ctypes.POINTER(_vp)]),
    "pm_batch_destroy": (None, [_vp]),
    "pm_batch_set_tuning": (_i32, [_vp, ctypes.POINTER(BatchTuning)]),
    "pm_batch_get_tuning": (_i32, [_vp, ctypes.POINTER(BatchTuning)]),
    "pm_batch_run": (_i32, [_vp, _vp]),
    "pm_batch_run_stage": (_i32, [_vp, _i32, _vp]),
    "pm_batch_autotune": (_i32, [_vp, _vp, _i32, ctypes.POINTER(BatchTuning)]),
    "pm_batch_capture": (_i32, [_vp, _vp]),
    "pm_batch_launch": (_i32, [_vp, _vp]),
    "pm_mapper_create": (_i32, [_i32, _i32, _i32, ctypes.POINTER(_vp)]),
    "pm_mapper_destroy": (None, [_vp]),
    "pm_mapper_init_intrinsic": (_i32, [_vp, _dp, ctypes.c_float, ctypes.c_float]),
    "pm_mapper_init_reference": (_i32, [_vp, _vp]),
    "pm_mapper_update": (_i32, [_vp, _vp, _dp, _dp, _dp, _dp]),
    "pm_mapper_init_reference_frame": (_i32, [_vp, _vp, ctypes.c_int64]),
    "pm_mapper_update_frame": (_i32, [_vp, _vp, ctypes.c_int64, _dp, _dp, _dp, _dp]),
    "pm_frame_ring_create": (_i32, [_i32, _i32, _i32, ctypes.POINTER(_vp)]),
    "pm_frame_ring_destroy": (None, [_vp]),
    "pm_frame_ring_put": (_i32, [_vp, ctypes.c_int64, _vp, _vp]),
    "pm_frame_ring_frame": (_i32, [_vp, ctypes.c_int64, ctypes.POINTER(_vp),
                                   ctypes.POINTER(_vp)]),
    "pm_frame_ring_puts": (_i32, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "pm_mapper_output": (_i32, [_vp, _vp]),
    "pm_mapper_output_fusion": (_i32, [_vp, _vp, _dp, _dp, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_float]),
    "pm_mapper_output_fusion_cpu": (_i32, [_vp, _vp, _dp, _dp, ctypes.c_float,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "pm_mapper_output_case": (_i32, [_vp, _i32, _vp]),
    "pm_mapper_download": (_i32, [_vp, ctypes.c_char_p, _vp, ctypes.c_size_t]),
    "pm_mapper_set_cost_type": (_i32, [_vp, _i32]),
    "pm_get_knobs": (None, [ctypes.POINTER(Knobs)]),
    "pm_set_knobs": (None, [ctypes.POINTER(Knobs)]),
    "pm_diff_create": (_i32, [Shape, _i32, SgmParams, DiffParams, ctypes.POINTER(_vp)]),
    "pm_diff_destroy": (None, [_vp]),
    "pm_diff_memory": (_i32, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "pm_diff_forward": (_i32, [_vp, _vp, ctypes.POINTER(_vp), ctypes.POINTER(Homography),
                               ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, _vp]),
    "pm_diff_backward": (_i32, [_vp, ctypes.c_float, _vp, _vp, _vp]),
    "pm_diff_download": (_i32, [_vp, ctypes.c_char_p, _vp, ctypes.c_size_t]),
    "pm_se3_homography": (_i32, [_dp, _dp, _dp, _dp, _dp, _dp]),
    "pm_se3_pose_grad": (_i32, [_dp, _dp, _dp, _dp, _dp, _dp]),
    "pm_depth_fuse_cpu_rules_workspace_size": (_i32, [_i32, _i32,
                                                      ctypes.POINTER(ctypes.c_uint64)]),
    "pm_depth_fuse_cpu_rules": (_i32, [_i32, _i32, Camera, _dp, _dp, _vp, _vp, _vp,
                                       ctypes.c_uint64, FusionState, FusionState, _vp]),
    "pm_driver_create": (_i32, [ctypes.POINTER(DriverConfig), _i32, ctypes.POINTER(_vp)]),
    "pm_driver_destroy": (None, [_vp]),
    "pm_driver_grab_image": (_i32, [_vp, ctypes.c_int64, _vp, _i32, _i32]),
    "pm_driver_grab_pose": (_i32, [_vp, ctypes.c_int64, _dp]),
    "pm_driver_process": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(_i32), ctypes.POINTER(_i32), _vp,
                                 ctypes.c_size_t]),
    "pm_driver_skipped": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "pm_driver_ring_puts": (_i32, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "pm_driver_buffered": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                                  ctypes.POINTER(_i32)]),
    "pm_quaternion_to_rotation": (_i32, [_dp, _dp]),
}

_lib = None


def load():
    """Load ``libpm_hip.so`` and bind every entry point (raises if the library is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m "
                              "probabilistic_mapping_amd.build` (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status: int) -> None:
    if status != PM_OK:
        raise PmError(status, load().pm_last_error().decode())


def last_error() -> str:
    return load().pm_last_error().decode()


def dptr(a) -> ctypes.POINTER(ctypes.c_double):
    import numpy as np
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)
