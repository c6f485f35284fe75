"""Oracle-side keyframe sequencing -- TEST INFRASTRUCTURE ONLY.

Restates the reference's StereoMapper sequencing (motion_stereo/stereo_mapper.cpp:63-244):
the FP64 homography setup rounded to float (:79-90), cost -> SGM -> filter per keyframe, and
the outputFusion loop -- the first keyframe only resets the state (:128-146), later ones fuse
previous -> current with R_p2c = R^T R_prev, t_p2c = R^T (t_prev - t) (:155-156) into
ping-pong state buffers (:160-243).  The homography is computed here with numpy, independently
of the library's C++ setup, so the two are checked against each other.
"""
from __future__ import annotations

import numpy as np

from . import calc_cost, depth_fuse, filter_cost, median3x3, sgm


def homography(K, R_l, T_l, R_r, T_r):
    """H = K R_r^T R_l K^-1, t = K R_r^T (T_l - T_r) (stereo_mapper.cpp:79-80), FP64, then
    rounded to float as the reference passes them (:82-90)."""
    K = np.asarray(K, np.float64)
    R_l = np.asarray(R_l, np.float64); R_r = np.asarray(R_r, np.float64)
    T_l = np.asarray(T_l, np.float64).reshape(3); T_r = np.asarray(T_r, np.float64).reshape(3)
    H = K @ R_r.T @ R_l @ np.linalg.inv(K)
    t = K @ R_r.T @ (T_l - T_r)
    return H.astype(np.float32), t.astype(np.float32)


def pose_p2c(R, t, R_prev, t_prev):
    R = np.asarray(R, np.float64); R_prev = np.asarray(R_prev, np.float64)
    t = np.asarray(t, np.float64).reshape(3); t_prev = np.asarray(t_prev, np.float64).reshape(3)
    return (R.T @ R_prev).astype(np.float32), (R.T @ (t_prev - t)).astype(np.float32)


class OracleSequence:
    """The keyframe loop of StereoMapper::outputFusion on the CPU oracle.

    ``precision`` = "f64" (the oracle) or "f32" (the CPU baseline).  After each keyframe the
    attributes ``cost``, ``agg``, ``depth_raw`` (filtered), ``depth`` (filtered, then the
    median if enabled: what fusion consumes), ``fused`` and ``state`` hold that keyframe's
    results.
    """

    def __init__(self, images, poses, sources, K, planes, idep_min, idep_max, knobs, paths,
                 median=False, precision="f64"):
        self.images = [np.ascontiguousarray(im, np.uint8) for im in images]
        self.poses = [(np.asarray(R, np.float64), np.asarray(T, np.float64).reshape(3))
                      for R, T in poses]
        self.sources = [list(s) for s in sources]
        self.K = np.asarray(K, np.float64)
        self.cam = (self.K[0, 0], self.K[1, 1], self.K[0, 2], self.K[1, 2])
        self.D = planes
        self.idep_min = float(idep_min)
        self.dep_sample = float((idep_max - idep_min) / (planes - 1))
        self.knobs = dict(knobs)
        self.paths = paths
        self.median = median
        self.precision = precision
        self.dt = np.float64 if precision == "f64" else np.float32
        H, W = self.images[0].shape
        self.states = [self._zero_state(H, W), self._zero_state(H, W)]
        self.first = True
        self.pre = 0
        self.prev_pose = None

    def _zero_state(self, H, W):
        return {k: np.zeros((H, W), self.dt) for k in ("alpha", "beta", "mu", "sig")}

    def depth_stage(self, i):
        srcs = self.sources[i]
        homs = [homography(self.K, *self.poses[i], *self.poses[j]) for j in srcs]
        self.cost = calc_cost(self.images[i], [self.images[j] for j in srcs], homs,
                              self.idep_min, self.dep_sample, self.D, self.precision)
        self.agg = sgm(self.images[i], self.cost, self.knobs, self.paths, self.precision)
        d = filter_cost(self.agg, self.knobs["var_scale"], self.idep_min, self.dep_sample,
                        self.precision)
        self.depth_raw = d
        self.depth = median3x3(d, self.precision) if self.median else d
        return self.depth

    def fusion_stage(self, i):
        R, t = self.poses[i]
        if self.first:
            self.first = False
            self.pre = 0
            for s in self.states:
                s["alpha"][:] = 0
            self.prev_pose = (R, t)
            self.fused = np.zeros_like(self.depth)
            self.table = None
            return
        self.fused = self.depth.copy()
        r, tt = pose_p2c(R, t, *self.prev_pose)
        self.table = depth_fuse(self.fused, self.states[self.pre], self.states[self.pre ^ 1],
                                r, tt, self.cam, self.precision)
        self.prev_pose = (R, t)
        self.pre ^= 1

    def run_keyframe(self, i):
        self.depth_stage(i)
        self.fusion_stage(i)

    @property
    def state(self):
        return self.states[self.pre]
