"""CPU oracle for the per-keyframe depth pipeline -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``probabilistic_mapping_amd``) never does.

``pm_oracle.hpp`` restates the reference's four stages in one templated C++ source:
``double`` is the FP64 oracle, ``float`` with OpenMP is the CPU baseline.  Citations of the
reference (HKUST-Aerial-Robotics/probabilistic_mapping) are in that header, one per stage:

* :func:`calc_cost`   -- motion_stereo/calc_cost.cu:11-210 (+ the running mean, :200-207)
* :func:`sgm`         -- motion_stereo/sgm.cu:5-135
* :func:`filter_cost` -- motion_stereo/filter_cost.cu:3-55
* :func:`depth_fuse`  -- motion_stereo/depth_fusion.cu:8-280
* :func:`median3x3`   -- not in the reference (optional 3x3 median of the 752x480 workload)

Pinning: every function is checked against hand-derived fixtures, closed forms and
independent brute-force formulations in ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PM_ORACLE_LIB selects another build of the same source (the -march=native one of
# ``make native``, used by the CPU baseline's timing process)
_LIB_PATH = os.environ.get("PM_ORACLE_LIB") or os.path.join(_HERE, "libpm_oracle.so")
_lib = None

_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build() -> str:
    """Compile ``libpm_oracle.so`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_max_threads.restype = ctypes.c_int
    return _lib


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def _ptr(a: np.ndarray, ctype):
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be C-contiguous"
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _dtype(precision: str):
    if precision == "f64":
        return np.float64, ctypes.c_double
    if precision == "f32":
        return np.float32, ctypes.c_float
    raise ValueError(precision)


def calc_cost(ref, srcs, homs, idep_min, dep_sample, planes, precision="f64", margin=False):
    """Fused-over-sources plane-sweep cost, pixel-major ``[H][W][D]``.

    ``srcs`` is a list of uint8 images, ``homs`` a list of ``(h[9], t[3])`` pairs holding the
    float-rounded homography setup of each measurement frame.  ``margin=True`` (FP64 only)
    also returns each entry's validity margin (float32 ``[H][W][D]``, pixels): the distance of
    the nearest source's decisive warp to the image-bounds test, i.e. how far the warps must
    move for the entry's set of contributing sources to change.
    """
    fref = np.asarray(ref).dtype == np.float64   # a real-valued reference (FD checks)
    ref = np.ascontiguousarray(ref, dtype=np.float64 if fref else np.uint8)
    H, W = ref.shape
    S = len(srcs)
    imgs = [np.ascontiguousarray(s, dtype=np.uint8) for s in srcs]
    for im in imgs:
        assert im.shape == (H, W)
    arr = (_u8p * max(S, 1))(*[_ptr(im, ctypes.c_uint8) for im in imgs])
    h = np.ascontiguousarray(np.array([np.asarray(a, np.float64).reshape(9) for a, _ in homs]
                                      if S else np.zeros((0, 9)), np.float64))
    t = np.ascontiguousarray(np.array([np.asarray(b, np.float64).reshape(3) for _, b in homs]
                                      if S else np.zeros((0, 3)), np.float64))
    dt, ct = _dtype(precision)
    out = np.empty((H, W, planes), dt)
    refp = _ptr(ref, ctypes.c_double if fref else ctypes.c_uint8)
    args = [ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(planes), refp,
            ctypes.c_int(S), arr, _ptr(h, ctypes.c_double), _ptr(t, ctypes.c_double),
            ctypes.c_double(idep_min), ctypes.c_double(dep_sample), _ptr(out, ct)]
    if fref:
        assert precision == "f64" and not margin, "a float64 reference runs the FP64 oracle"
        lib().oracle_calc_cost_fref_f64(*args)
        return out
    if margin:
        assert precision == "f64", "the validity margin comes with the FP64 oracle"
        mg = np.empty((H, W, planes), np.float32)
        lib().oracle_calc_cost_margin_f64(*args, _ptr(mg, ctypes.c_float))
        return out, mg
    fn = lib().oracle_calc_cost_f64 if precision == "f64" else lib().oracle_calc_cost_f32
    fn(*args)
    return out


def _knob_array(knobs) -> np.ndarray:
    return np.array([knobs["pi1"], knobs["pi2"], knobs["tau_so"], knobs["sgm_q1"],
                     knobs["sgm_q2"], knobs["var_scale"]], np.float64)


def sgm(ref, cost, knobs, paths, precision="f64"):
    """Multi-path semi-global aggregation of a ``[H][W][D]`` cost volume.  A float64
    ``ref`` (FP64 only) is used as real-valued intensities for the P1/P2 switch."""
    fref = np.asarray(ref).dtype == np.float64
    ref = np.ascontiguousarray(ref, dtype=np.float64 if fref else np.uint8)
    H, W = ref.shape
    dt, ct = _dtype(precision)
    cost = np.ascontiguousarray(cost, dtype=dt)
    D = cost.shape[2]
    agg = np.empty_like(cost)
    kn = _knob_array(knobs)
    if fref:
        assert precision == "f64", "a float64 reference runs the FP64 oracle"
        fn, refp = lib().oracle_sgm_fref_f64, _ptr(ref, ctypes.c_double)
    else:
        fn = lib().oracle_sgm_f64 if precision == "f64" else lib().oracle_sgm_f32
        refp = _ptr(ref, ctypes.c_uint8)
    fn(ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(D), refp,
       _ptr(cost, ct), _ptr(kn, ctypes.c_double), ctypes.c_int(paths), _ptr(agg, ct))
    return agg


def sgm_adjoint(ref, cost, knobs, paths, gagg, precision="f64"):
    """dL/dC of the multi-path aggregation given dL/dagg (both ``[H][W][D]``, float64 out).
    ``precision`` sets the arithmetic of the recomputed forward and of its routing decisions
    ("f32" reproduces an FP32 forward); gradients accumulate in double."""
    ref = np.ascontiguousarray(ref, dtype=np.uint8)
    H, W = ref.shape
    dt, ct = _dtype(precision)
    cost = np.ascontiguousarray(cost, dtype=dt)
    D = cost.shape[2]
    g = np.ascontiguousarray(gagg, dtype=np.float64)
    assert g.shape == cost.shape
    out = np.empty(cost.shape, np.float64)
    kn = _knob_array(dict(knobs, var_scale=knobs.get("var_scale", 1.0)))
    fn = lib().oracle_sgm_adjoint_f64 if precision == "f64" else lib().oracle_sgm_adjoint_f32
    fn(ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(D), _ptr(ref, ctypes.c_uint8),
       _ptr(cost, ct), _ptr(kn, ctypes.c_double), ctypes.c_int(paths),
       _ptr(g, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out


def filter_cost(agg, var_scale, idep_min, dep_sample, precision="f64"):
    """WTA + rejection rules + parabola subpixel fit -> depth map ``[H][W]``."""
    dt, ct = _dtype(precision)
    agg = np.ascontiguousarray(agg, dtype=dt)
    H, W, D = agg.shape
    depth = np.empty((H, W), dt)
    fn = lib().oracle_filter_f64 if precision == "f64" else lib().oracle_filter_f32
    fn(ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(D), _ptr(agg, ct),
       ctypes.c_double(var_scale), ctypes.c_double(idep_min), ctypes.c_double(dep_sample),
       _ptr(depth, ct))
    return depth


def median3x3(depth, precision="f64"):
    dt, ct = _dtype(precision)
    depth = np.ascontiguousarray(depth, dtype=dt)
    H, W = depth.shape
    out = np.empty_like(depth)
    fn = lib().oracle_median_f64 if precision == "f64" else lib().oracle_median_f32
    fn(ctypes.c_int(W), ctypes.c_int(H), _ptr(depth, ct), _ptr(out, ct))
    return out


def depth_fuse_cpu_rules(depth, pre, cur, r_p2c, t_p2c, intr, precision="f64"):
    """The fusion step of outputFusionCPU with CASE >= 3 (stereo_mapper.cpp:719-800), in place
    like :func:`depth_fuse`: ``depth`` is the filtered depth in and the result out.  Returns
    the filled propagation table (int32, (y << 16) | x of the source, 0 = empty)."""
    return _fuse(depth, pre, cur, r_p2c, t_p2c, intr, precision, "oracle_fuse_cpu_rules")


def depth_fuse(depth, pre, cur, r_p2c, t_p2c, intr, precision="f64"):
    """One Gaussian x Beta fusion step, in place on ``depth``, ``pre`` and ``cur``.

    ``pre`` / ``cur`` are dicts of ``alpha, beta, mu, sig`` ``[H][W]`` arrays of the chosen
    precision.  Returns the filled propagation table (int32, 0 = empty).
    """
    return _fuse(depth, pre, cur, r_p2c, t_p2c, intr, precision, "oracle_fuse")


def _fuse(depth, pre, cur, r_p2c, t_p2c, intr, precision, name):
    dt, ct = _dtype(precision)
    H, W = depth.shape
    for a in [depth] + [pre[k] for k in "alpha beta mu sig".split()] + \
             [cur[k] for k in "alpha beta mu sig".split()]:
        assert a.dtype == dt and a.flags["C_CONTIGUOUS"] and a.shape == (H, W)
    table = np.zeros((H, W), np.int32)
    P = ctypes.POINTER(ct)
    pre_arr = (P * 4)(*[_ptr(pre[k], ct) for k in ("alpha", "beta", "mu", "sig")])
    cur_arr = (P * 4)(*[_ptr(cur[k], ct) for k in ("alpha", "beta", "mu", "sig")])
    intr = np.ascontiguousarray(np.asarray(intr, np.float64).reshape(4))
    r = np.ascontiguousarray(np.asarray(r_p2c, np.float64).reshape(9))
    t = np.ascontiguousarray(np.asarray(t_p2c, np.float64).reshape(3))
    fn = getattr(lib(), f"{name}_{precision}")
    fn(ctypes.c_int(W), ctypes.c_int(H), _ptr(intr, ctypes.c_double),
       _ptr(r, ctypes.c_double), _ptr(t, ctypes.c_double), _ptr(depth, ct),
       _ptr(table, ctypes.c_int32), pre_arr, cur_arr)
    return table
