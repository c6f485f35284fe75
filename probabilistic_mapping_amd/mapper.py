"""Python mirror of the reference's ``StereoMapper`` (motion_stereo/stereo_mapper.h:26-63).

Same method names, argument order and meaning; numpy arrays stand in for cv::Mat / Eigen.
Calls go through the C ABI (``pm_mapper_*``) to the C++ class in ``csrc/stereo_mapper.cpp``.
Errors raise :class:`PmError` instead of terminating the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C

WIDTH, HEIGHT, DEP_CNT = 640, 480, 64   # motion_stereo/parameters.h:8-15


def _d(a, n):
    a = np.ascontiguousarray(np.asarray(a, np.float64).reshape(n))
    return a, C.dptr(a)


def get_knobs() -> dict:
    k = C.Knobs()
    C.load().pm_get_knobs(ctypes.byref(k))
    return {f: getattr(k, f) for f, _ in C.Knobs._fields_}


def set_knobs(**kw) -> dict:
    """Sets the library's knob globals (pi1, pi2, tau_so, sgm_q1, sgm_q2, var_scale,
    case_mode, idep_min, sgm_paths, median); returns the previous values."""
    old = get_knobs()
    k = C.Knobs()
    for f, _ in C.Knobs._fields_:
        setattr(k, f, kw.get(f, old[f]))
    unknown = set(kw) - {f for f, _ in C.Knobs._fields_}
    if unknown:
        raise KeyError(f"unknown knobs: {sorted(unknown)}")
    C.load().pm_set_knobs(ctypes.byref(k))
    return old


class FrameRing:
    """Upload-once frame ring (``pm_frame_ring_*``): frame ``f`` is uploaded and converted
    once and stays on the device in slot ``f % capacity`` until another frame takes it."""

    def __init__(self, width: int, height: int, capacity: int):
        self._lib = C.load()
        h = ctypes.c_void_p()
        C.check(self._lib.pm_frame_ring_create(width, height, capacity, ctypes.byref(h)))
        self._h = h
        self.width, self.height, self.capacity = width, height, capacity

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pm_frame_ring_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def put(self, frame: int, image, stream=None):
        img = np.ascontiguousarray(image)
        if img.dtype != np.uint8 or img.shape != (self.height, self.width):
            raise C.PmError(C.PM_ERR_INVALID_ARGUMENT,
                            f"image must be uint8 {self.height}x{self.width}")
        C.check(self._lib.pm_frame_ring_put(self._h, int(frame),
                                            img.ctypes.data_as(ctypes.c_void_p), stream))

    def device_pointers(self, frame: int) -> tuple[int, int]:
        """(u8 image, quad texture) device addresses of a held frame."""
        img, quad = ctypes.c_void_p(), ctypes.c_void_p()
        C.check(self._lib.pm_frame_ring_frame(self._h, int(frame), ctypes.byref(img),
                                              ctypes.byref(quad)))
        return img.value, quad.value

    def holds(self, frame: int) -> bool:
        try:
            self.device_pointers(frame)
            return True
        except C.PmError:
            return False

    @property
    def puts(self) -> int:
        n = ctypes.c_int64(0)
        C.check(self._lib.pm_frame_ring_puts(self._h, ctypes.byref(n)))
        return n.value


class StereoMapper:
    def __init__(self, width: int = WIDTH, height: int = HEIGHT, dep_cnt: int = DEP_CNT):
        self._lib = C.load()
        h = ctypes.c_void_p()
        C.check(self._lib.pm_mapper_create(width, height, dep_cnt, ctypes.byref(h)))
        self._h = h
        self.width, self.height, self.dep_cnt = width, height, dep_cnt

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pm_mapper_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_cost_type(self, cost_type: str):
        code = {"f32": C.PM_COST_F32, "f16": C.PM_COST_F16, "bf16": C.PM_COST_BF16}[cost_type]
        C.check(self._lib.pm_mapper_set_cost_type(self._h, code))

    def initIntrinsic(self, K, bf: float, dep_sample: float):
        k, kp = _d(K, 9)
        C.check(self._lib.pm_mapper_init_intrinsic(self._h, kp, float(bf), float(dep_sample)))

    def _img(self, img):
        img = np.ascontiguousarray(img)
        if img.dtype != np.uint8 or img.shape != (self.height, self.width):
            raise C.PmError(C.PM_ERR_INVALID_ARGUMENT,
                            f"image must be uint8 {self.height}x{self.width}")
        return img

    def initReference(self, img_l):
        img = self._img(img_l)
        C.check(self._lib.pm_mapper_init_reference(self._h, img.ctypes.data_as(ctypes.c_void_p)))

    def update(self, img_r, R_l, T_l, R_r, T_r):
        img = self._img(img_r)
        a = [_d(R_l, 9), _d(T_l, 3), _d(R_r, 9), _d(T_r, 3)]
        C.check(self._lib.pm_mapper_update(self._h, img.ctypes.data_as(ctypes.c_void_p),
                                           *[p for _, p in a]))

    def initReferenceFrame(self, ring: FrameRing, frame: int):
        """initReference with frame ``frame`` of ``ring`` (no upload)."""
        C.check(self._lib.pm_mapper_init_reference_frame(self._h, ring._h, int(frame)))

    def updateFrame(self, ring: FrameRing, frame: int, R_l, T_l, R_r, T_r):
        """update with frame ``frame`` of ``ring``; it must stay in the ring until the next
        output call returns."""
        a = [_d(R_l, 9), _d(T_l, 3), _d(R_r, 9), _d(T_r, 3)]
        C.check(self._lib.pm_mapper_update_frame(self._h, ring._h, int(frame),
                                                 *[p for _, p in a]))

    def _result(self):
        return np.empty((self.height, self.width), np.float32)

    def output(self) -> np.ndarray:
        r = self._result()
        C.check(self._lib.pm_mapper_output(self._h, r.ctypes.data_as(ctypes.c_void_p)))
        return r

    def outputFusion(self, R, t, fx, fy, cx, cy) -> np.ndarray:
        r = self._result()
        (Ra, Rp), (ta, tp) = _d(R, 9), _d(t, 3)
        C.check(self._lib.pm_mapper_output_fusion(self._h, r.ctypes.data_as(ctypes.c_void_p),
                                                  Rp, tp, float(fx), float(fy), float(cx),
                                                  float(cy)))
        return r

    def outputFusionCPU(self, R, t, fx, fy, cx, cy) -> np.ndarray:
        """Dispatches on the CASE knob (``case_mode``): 0-2 give WTA / filtered depths, 3 and
        above the fused depth of the reference's CPU fusion rules."""
        r = self._result()
        (Ra, Rp), (ta, tp) = _d(R, 9), _d(t, 3)
        C.check(self._lib.pm_mapper_output_fusion_cpu(self._h, r.ctypes.data_as(ctypes.c_void_p),
                                                      Rp, tp, float(fx), float(fy), float(cx),
                                                      float(cy)))
        return r

    def download(self, name: str) -> np.ndarray:
        H, W, D = self.height, self.width, self.dep_cnt
        shapes = {"raw_cost": ((H, W, D), np.float32), "sgm_cost": ((H, W, D), np.float32),
                  "table": ((H, W), np.int32), "cpu_table": ((H, W), np.int32),
                  "img_l": ((H, W), np.uint8)}
        shape, dt = shapes.get(name, ((H, W), np.float32))
        out = np.empty(shape, dt)
        C.check(self._lib.pm_mapper_download(self._h, name.encode(),
                                             out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out
