"""Keyframe / source selection driver: the reference node's processImg without ROS.

``FrameDriver`` wraps ``pm_driver_*`` (the C++ ``pm::FrameDriver``, include/pm/frame_driver.h;
reference: evaluate.cc:75-110 GrabImage / GrabPose, :137-276 processImg).  The dataset helpers
write and read the files of the ``pm_evaluate`` executable: 8-bit binary PGM frames listed as
"timestamp path" lines, poses as "timestamp tx ty tz qx qy qz qw" lines (the TUM RGB-D
layouts), and float32 ``.npy`` depth maps.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import _capi as C

EVALUATE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pm_evaluate")


def quaternion_to_rotation(q) -> np.ndarray:
    """Eigen's ``Quaterniond(w, x, y, z).toRotationMatrix()`` for q = (x, y, z, w)."""
    qa = np.ascontiguousarray(q, np.float64)
    R = np.empty(9, np.float64)
    C.check(C.load().pm_quaternion_to_rotation(C.dptr(qa), C.dptr(R)))
    return R.reshape(3, 3)


class FrameDriver:
    """``with_mapper=False`` plans the steps only (host code, no GPU)."""

    def __init__(self, agg_num=3, interval=6, planes=64, fx=481.2017, fy=-480.0002, cx=319.5,
                 cy=239.5, dep_sample=1.0 / (0.15 * 160.0), with_mapper=True, frame_ring=True):
        self.lib = C.load()
        self.agg_num = agg_num
        cfg = C.DriverConfig(agg_num, interval, planes, fx, fy, cx, cy, dep_sample,
                             1 if frame_ring else 0)
        h = ctypes.c_void_p()
        C.check(self.lib.pm_driver_create(ctypes.byref(cfg), 1 if with_mapper else 0,
                                          ctypes.byref(h)))
        self._h = h
        self.with_mapper = with_mapper
        self.shape = None

    def close(self):
        if self._h:
            self.lib.pm_driver_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def grab_image(self, stamp_ns: int, image: np.ndarray):
        img = np.ascontiguousarray(image, np.uint8)
        self.shape = img.shape
        C.check(self.lib.pm_driver_grab_image(self._h, int(stamp_ns), img.ctypes.data,
                                              img.shape[1], img.shape[0]))

    def grab_pose(self, stamp_ns: int, t, q):
        """t = (tx, ty, tz), q = (qx, qy, qz, qw): camera -> world."""
        p = np.ascontiguousarray([*t, *q], np.float64)
        C.check(self.lib.pm_driver_grab_pose(self._h, int(stamp_ns), C.dptr(p)))

    def process(self):
        """One processImg step: None, or (stamp_ns, ref_image, [src_images], depth or None)."""
        ran = ctypes.c_int32(0)
        stamp = ctypes.c_int64(0)
        ref = ctypes.c_int32(-1)
        srcs = (ctypes.c_int32 * max(self.agg_num, 1))()
        depth = None
        if self.with_mapper and self.shape is not None:
            depth = np.empty(self.shape, np.float32)
        C.check(self.lib.pm_driver_process(
            self._h, ctypes.byref(ran), ctypes.byref(stamp), ctypes.byref(ref), srcs,
            depth.ctypes.data if depth is not None else None,
            depth.nbytes if depth is not None else 0))
        if not ran.value:
            return None
        return stamp.value, ref.value, list(srcs)[:self.agg_num], depth

    def buffered(self) -> tuple[int, int, int]:
        """(images, poses, images still holding pixels) in the message lists."""
        a, b, c = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        C.check(self.lib.pm_driver_buffered(self._h, ctypes.byref(a), ctypes.byref(b),
                                            ctypes.byref(c)))
        return a.value, b.value, c.value

    @property
    def ring_puts(self) -> int:
        """Frames uploaded through the frame ring so far (0 without one)."""
        n = ctypes.c_int64(0)
        C.check(self.lib.pm_driver_ring_puts(self._h, ctypes.byref(n)))
        return n.value

    @property
    def skipped(self) -> int:
        n = ctypes.c_int32(0)
        C.check(self.lib.pm_driver_skipped(self._h, ctypes.byref(n)))
        return n.value


# ---- dataset files of pm_evaluate ---------------------------------------------------------

def stamp_text(ns: int) -> str:
    return f"{ns // 1_000_000_000}.{ns % 1_000_000_000:09d}"


def write_pgm(path: str, image: np.ndarray) -> None:
    img = np.ascontiguousarray(image, np.uint8)
    with open(path, "wb") as f:
        f.write(f"P5\n# frame\n{img.shape[1]} {img.shape[0]}\n255\n".encode())
        f.write(img.tobytes())


def write_dataset(root: str, stamps, images, poses) -> tuple[str, str]:
    """stamps: integer nanoseconds; poses: (t[3], q[4]) per stamp.  Returns the two list files."""
    os.makedirs(os.path.join(root, "frames"), exist_ok=True)
    img_list = os.path.join(root, "images.txt")
    pose_list = os.path.join(root, "poses.txt")
    with open(img_list, "w") as fi, open(pose_list, "w") as fp:
        fi.write("# timestamp filename\n")
        fp.write("# timestamp tx ty tz qx qy qz qw\n")
        for ns, img, (t, q) in zip(stamps, images, poses):
            name = f"frames/{stamp_text(ns)}.pgm"
            write_pgm(os.path.join(root, name), img)
            fi.write(f"{stamp_text(ns)} {name}\n")
            fp.write(f"{stamp_text(ns)} " + " ".join(repr(float(v)) for v in (*t, *q)) + "\n")
    return img_list, pose_list


def run_evaluate(img_list: str, pose_list: str, out_dir: str, case=None, plan_only=False,
                 frame_ring=True, timeout=600, **cfg) -> list:
    """Runs the executable; returns its steps as (stamp text, ref, [srcs], depth or None)."""
    args = [EVALUATE, img_list, pose_list, out_dir]
    if case is not None:
        args.append(str(case))
    for k, v in cfg.items():
        args += ["--" + k.replace("_", "-"), repr(v)]
    if plan_only:
        args.append("--plan-only")
    if not frame_ring:
        args.append("--no-frame-ring")
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"pm_evaluate failed ({r.returncode}): {r.stderr}")
    steps = []
    if plan_only:
        for line in r.stdout.splitlines():
            f = line.split()
            steps.append((f[0], int(f[1]), [int(x) for x in f[2:]], None))
        return steps
    with open(os.path.join(out_dir, "depths.txt")) as fl:
        for line in fl:
            f = line.split()
            steps.append((f[0], int(f[2]), [int(x) for x in f[3:]],
                          np.load(os.path.join(out_dir, f[1]))))
    return steps
