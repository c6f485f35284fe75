"""Seeded synthetic scenes for the depth pipeline.

Scenes are textured rectangles (free planes and box faces) with multi-octave value-noise
albedo in 0..255, rendered by exact ray/plane intersection in ``csrc/synth.cpp``.  Poses
follow the reference's ``StereoMapper::update`` convention (motion_stereo/stereo_mapper.cpp:
73-80): ``R`` is the camera->world rotation and ``T`` the camera centre in world, so a world
point is ``X_w = R X_c + T``.

Workloads:

* :func:`small_scene`   -- 320x240, 1 reference + 2 measurement frames, 4 textured planes at
  1-5 m, fx=fy=250, 32 planes over inverse depth 0.2-1.0, 4-path SGM (P1=0.003, P2=0.01).
* :func:`trajectory_scene` -- 752x480, 12 random planes and boxes at 0.5-8 m, fx=fy=460,
  a 20-frame sinusoidal trajectory with 3-15 cm spacing, 5 nearest frames per reference,
  64 planes over inverse depth 0.125-2.0, 8-path SGM.
* :func:`uneven_keyframes` -- 16 independent keyframes, each from its own scene and
  trajectory, with the resolution drawn from {320x240, 640x480, 752x480, 1280x720}, the
  source count from 2..8 and the plane count from {32, 64, 96, 128}.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SYNTH_LIB = os.path.join(_HERE, "libpm_synth.so")
_lib = None


def _synth_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SYNTH_LIB):
            from . import build as _build
            _build.build_synth()
        _lib = ctypes.CDLL(_SYNTH_LIB)
        _lib.pm_synth_render.restype = ctypes.c_long
    return _lib


def rot_xyz(rx: float, ry: float, rz: float) -> np.ndarray:
    """Rotation matrix Rz @ Ry @ Rx from angles in radians."""
    cx, sx, cy, sy, cz, sz = (math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry),
                              math.cos(rz), math.sin(rz))
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def axis_angle(axis, angle: float) -> np.ndarray:
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    Kx = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * Kx + (1 - math.cos(angle)) * (Kx @ Kx)


@dataclass
class Scene:
    width: int
    height: int
    K: np.ndarray                       # 3x3 intrinsics
    patches: np.ndarray                 # [n, 16] patch records (see csrc/synth.cpp)
    poses: list = field(default_factory=list)   # [(R 3x3, T 3)] camera->world

    @property
    def intrinsics(self):
        return (float(self.K[0, 0]), float(self.K[1, 1]), float(self.K[0, 2]),
                float(self.K[1, 2]))

    def render(self, R, T):
        """Returns (uint8 image [H][W], float32 camera-z depth [H][W])."""
        img = np.zeros((self.height, self.width), np.uint8)
        dep = np.zeros((self.height, self.width), np.float32)
        p = np.ascontiguousarray(self.patches, np.float64)
        k = np.array(self.intrinsics, np.float64)
        R = np.ascontiguousarray(R, np.float64)
        T = np.ascontiguousarray(np.asarray(T, np.float64).reshape(3))
        dp = ctypes.POINTER(ctypes.c_double)
        _synth_lib().pm_synth_render(
            p.ctypes.data_as(dp), ctypes.c_int(len(p)), k.ctypes.data_as(dp),
            R.ctypes.data_as(dp), T.ctypes.data_as(dp), ctypes.c_int(self.width),
            ctypes.c_int(self.height), img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            dep.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return img, dep


def _patch(c, e1, e2, h1, h2, seed, freq, octaves=4, mean=128.0, amp=150.0):
    e1 = np.asarray(e1, np.float64); e1 = e1 / np.linalg.norm(e1)
    e2 = np.asarray(e2, np.float64)
    e2 = e2 - e1 * float(e1 @ e2); e2 = e2 / np.linalg.norm(e2)
    return np.concatenate([np.asarray(c, np.float64), e1, e2,
                           [h1, h2, float(seed), freq, float(octaves), mean, amp]])


def _box(rng, centre, size, yaw, seed0):
    """Six faces of a box with half sizes ``size`` rotated by ``yaw`` about world y."""
    R = rot_xyz(0.0, yaw, 0.0)
    out = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            n = R[:, axis] * sign
            a1, a2 = R[:, (axis + 1) % 3], R[:, (axis + 2) % 3]
            c = np.asarray(centre) + n * size[axis]
            out.append(_patch(c, a1, a2, size[(axis + 1) % 3], size[(axis + 2) % 3],
                              seed0 + len(out), freq=float(rng.uniform(3.0, 6.0)),
                              mean=float(rng.uniform(90, 170)), amp=float(rng.uniform(110, 170))))
    return out


def small_scene(seed: int = 0):
    """320x240 reference + 2 measurement frames, 4 textured planes at 1-5 m.

    Returns a dict with the scene, reference/measurement images and poses, ground-truth
    depth and the pipeline parameters of this workload.
    """
    rng = np.random.default_rng(seed)
    W, H = 320, 240
    K = np.array([[250.0, 0, 160.0], [0, 250.0, 120.0], [0, 0, 1.0]])
    patches = []
    # background wall near 4.6 m, slightly tilted, unbounded
    n = rot_xyz(rng.uniform(-0.1, 0.1), rng.uniform(-0.15, 0.15), 0.0)
    patches.append(_patch([0, 0, rng.uniform(4.3, 4.8)], n[:, 0], n[:, 1], 0, 0,
                          seed * 101 + 1, freq=2.5, amp=160))
    # three finite planes at 1.2-3.8 m
    for k in range(3):
        z = rng.uniform(1.2, 3.8)
        cx_ = rng.uniform(-0.35, 0.35) * z
        cy_ = rng.uniform(-0.25, 0.25) * z
        Rk = rot_xyz(rng.uniform(-0.5, 0.5), rng.uniform(-0.6, 0.6), rng.uniform(-0.3, 0.3))
        patches.append(_patch([cx_, cy_, z], Rk[:, 0], Rk[:, 1],
                              rng.uniform(0.15, 0.35) * z, rng.uniform(0.12, 0.3) * z,
                              seed * 101 + 2 + k, freq=float(rng.uniform(4, 8)),
                              mean=float(rng.uniform(100, 160)), amp=160))
    scene = Scene(W, H, K, np.array(patches))
    poses = [(np.eye(3), np.zeros(3))]
    for _ in range(2):
        d = rng.normal(size=3); d[2] *= 0.3; d /= np.linalg.norm(d)
        T = d * rng.uniform(0.05, 0.10)
        R = axis_angle(rng.normal(size=3), math.radians(rng.uniform(0.3, 2.0)))
        poses.append((R, T))
    scene.poses = poses
    frames = [scene.render(R, T) for R, T in poses]
    return {
        "scene": scene,
        "ref": frames[0][0], "ref_pose": poses[0], "gt_depth": frames[0][1],
        "srcs": [f[0] for f in frames[1:]], "src_poses": poses[1:],
        "planes": 32, "idep_min": 0.2, "idep_max": 1.0, "paths": 4,
        "knobs": dict(pi1=0.003, pi2=0.01, tau_so=4.0, sgm_q1=2.0, sgm_q2=3.0, var_scale=1.02),
    }


def trajectory_poses(n_frames: int, rng) -> list:
    """Sinusoidal camera path: lateral motion with 3-15 cm steps plus small wobble."""
    phase = rng.uniform(0, 2 * math.pi)
    steps = [0.09 + 0.06 * math.sin(2 * math.pi * i / max(n_frames - 1, 1) + phase)
             for i in range(n_frames)]
    x = np.concatenate([[0.0], np.cumsum(steps[:-1])])
    poses = []
    for i in range(n_frames):
        T = np.array([x[i] - x[-1] / 2, 0.03 * math.sin(2 * math.pi * i / 9.0),
                      0.03 * math.sin(2 * math.pi * i / 7.0 + 1.0)])
        R = rot_xyz(math.radians(1.5) * math.sin(2 * math.pi * i / 13.0),
                    math.radians(3.0) * math.sin(2 * math.pi * i / 20.0 + phase),
                    math.radians(1.0) * math.sin(2 * math.pi * i / 11.0))
        poses.append((R, T))
    return poses


def trajectory_scene(seed: int = 0, width: int = 752, height: int = 480, n_frames: int = 20,
                     n_sources: int = 5, n_objects: int = 12, planes: int = 64,
                     fx: float = 460.0):
    """752x480 sequence: 12 random planes/boxes at 0.5-8 m, 20-frame trajectory.

    Returns the scene, all frames with poses and ground truth, and for every keyframe the
    indices of its ``n_sources`` nearest frames (by camera-centre distance).
    """
    rng = np.random.default_rng(1000 + seed)
    W, H = width, height
    K = np.array([[fx, 0, W / 2.0], [0, fx, H / 2.0], [0, 0, 1.0]])
    patches = []
    n = rot_xyz(rng.uniform(-0.08, 0.08), rng.uniform(-0.1, 0.1), 0.0)
    patches.append(_patch([0, 0, 7.6], n[:, 0], n[:, 1], 0, 0, seed * 997 + 1, freq=1.5,
                          amp=160))
    for k in range(n_objects):
        z = float(np.exp(rng.uniform(np.log(0.7), np.log(6.5))))
        cx_ = rng.uniform(-0.45, 0.45) * z * W / (2 * fx) * 2
        cy_ = rng.uniform(-0.4, 0.4) * z * H / (2 * fx) * 2
        if k % 3 == 2:
            s = rng.uniform(0.08, 0.22) * z
            patches.extend(_box(rng, [cx_, cy_, z], [s, s * rng.uniform(0.6, 1.4), s],
                                rng.uniform(-0.8, 0.8), seed * 997 + 100 * (k + 1)))
        else:
            Rk = rot_xyz(rng.uniform(-0.6, 0.6), rng.uniform(-0.7, 0.7), rng.uniform(-0.4, 0.4))
            patches.append(_patch([cx_, cy_, z], Rk[:, 0], Rk[:, 1],
                                  rng.uniform(0.1, 0.3) * z, rng.uniform(0.08, 0.25) * z,
                                  seed * 997 + 10 + k, freq=float(rng.uniform(2.5, 6.0)),
                                  mean=float(rng.uniform(90, 170)), amp=float(rng.uniform(120, 170))))
    scene = Scene(W, H, K, np.array(patches))
    scene.poses = trajectory_poses(n_frames, rng)
    frames = [scene.render(R, T) for R, T in scene.poses]
    centres = np.array([T for _, T in scene.poses])
    sources = []
    for i in range(n_frames):
        dist = np.linalg.norm(centres - centres[i], axis=1)
        order = [j for j in np.argsort(dist, kind="stable") if j != i]
        sources.append([int(j) for j in order[:n_sources]])
    return {
        "scene": scene,
        "images": [f[0] for f in frames], "gt_depths": [f[1] for f in frames],
        "poses": scene.poses, "sources": sources,
        "planes": planes, "idep_min": 0.125, "idep_max": 2.0, "paths": 8,
        "knobs": dict(pi1=8.0, pi2=24.0, tau_so=4.0, sgm_q1=2.0, sgm_q2=3.0, var_scale=1.02),
    }


def diff_scene(seed: int = 0, width: int = 640, height: int = 480, n_sources: int = 4,
               planes: int = 64):
    """Reference frame + ``n_sources`` measurement frames of 4 textured planes at 1-5 m, with
    ground-truth depth, for the differentiable readout (focal length 250 px per 320 px of width).

    Poses are also given relative to the reference camera: (R, t) with
    X_src = R X_ref + t, i.e. R = R_r^T R_l, t = R_r^T (T_l - T_r) (stereo_mapper.cpp:79-80).
    """
    rng = np.random.default_rng(1000 + seed)
    f = 250.0 * width / 320.0
    K = np.array([[f, 0, width / 2.0], [0, f, height / 2.0], [0, 0, 1.0]])
    patches = []
    n = rot_xyz(rng.uniform(-0.1, 0.1), rng.uniform(-0.15, 0.15), 0.0)
    patches.append(_patch([0, 0, rng.uniform(4.3, 4.8)], n[:, 0], n[:, 1], 0, 0,
                          seed * 101 + 51, freq=2.5, amp=160))
    for k in range(3):
        z = rng.uniform(1.2, 3.8)
        cx_ = rng.uniform(-0.35, 0.35) * z
        cy_ = rng.uniform(-0.25, 0.25) * z
        Rk = rot_xyz(rng.uniform(-0.5, 0.5), rng.uniform(-0.6, 0.6), rng.uniform(-0.3, 0.3))
        patches.append(_patch([cx_, cy_, z], Rk[:, 0], Rk[:, 1],
                              rng.uniform(0.15, 0.35) * z, rng.uniform(0.12, 0.3) * z,
                              seed * 101 + 52 + k, freq=float(rng.uniform(4, 8)),
                              mean=float(rng.uniform(100, 160)), amp=160))
    scene = Scene(width, height, K, np.array(patches))
    poses = [(np.eye(3), np.zeros(3))]
    for _ in range(n_sources):
        d = rng.normal(size=3); d[2] *= 0.3; d /= np.linalg.norm(d)
        T = d * rng.uniform(0.05, 0.10)
        R = axis_angle(rng.normal(size=3), math.radians(rng.uniform(0.3, 2.0)))
        poses.append((R, T))
    scene.poses = poses
    frames = [scene.render(R, T) for R, T in poses]
    R_l, T_l = poses[0]
    rel = [(R_r.T @ R_l, R_r.T @ (T_l - T_r)) for R_r, T_r in poses[1:]]
    return {
        "scene": scene, "K": K,
        "ref": frames[0][0], "gt_depth": frames[0][1],
        "srcs": [fr[0] for fr in frames[1:]], "rel_poses": rel,
        "planes": planes, "idep_min": 0.2, "idep_max": 1.0, "paths": 8,
        "knobs": dict(pi1=8.0, pi2=24.0, tau_so=4.0, sgm_q1=2.0, sgm_q2=3.0),
        "temperature": 0.05, "cost_scale": 255.0,
    }


UNEVEN_RESOLUTIONS = ((320, 240), (640, 480), (752, 480), (1280, 720))
UNEVEN_PLANES = (32, 64, 96, 128)
# seed of the default 16-keyframe draw: it holds all four resolutions and plane counts, source
# counts 2..8, and the largest keyframe shape (1280x720, 128 planes, 8 sources)
UNEVEN_SEED = 4


def uneven_keyframes(seed: int = UNEVEN_SEED, count: int = 16):
    """``count`` independent keyframes of uneven shape.  Per keyframe the seeded draw picks a
    resolution from ``UNEVEN_RESOLUTIONS``, a source count uniform in 2..8 and a plane count
    from ``UNEVEN_PLANES``; the keyframe is the middle frame of its own seeded scene's
    ``S + 1``-frame trajectory (fx = fy = 0.612 W, the 752x480 workload's ratio) and its
    sources are the other ``S`` frames.  Each entry holds the frames, poses, intrinsics,
    ground-truth depth of the reference and the pipeline parameters."""
    rng = np.random.default_rng(5000 + seed)
    out = []
    for k in range(count):
        W, H = UNEVEN_RESOLUTIONS[int(rng.integers(len(UNEVEN_RESOLUTIONS)))]
        S = int(rng.integers(2, 9))
        D = int(rng.choice(UNEVEN_PLANES))
        sc = trajectory_scene(100 * seed + k + 50, width=W, height=H, n_frames=S + 1,
                              n_sources=S, planes=D, fx=0.612 * W)
        ref = S // 2
        srcs = [j for j in range(S + 1) if j != ref]
        out.append({
            "width": W, "height": H, "planes": D, "num_sources": S,
            "K": sc["scene"].K, "ref": sc["images"][ref], "ref_pose": sc["poses"][ref],
            "gt_depth": sc["gt_depths"][ref], "srcs": [sc["images"][j] for j in srcs],
            "src_poses": [sc["poses"][j] for j in srcs],
            "idep_min": sc["idep_min"], "idep_max": sc["idep_max"], "paths": sc["paths"],
            "knobs": sc["knobs"],
        })
    return out
