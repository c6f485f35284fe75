"""Differentiable soft-argmin depth readout over the SGM aggregate (C ABI), and the
``torch.autograd.Function`` that exposes it with respect to the se3 pose of every measurement
frame and the reference intensities.

The forward and the hand-written adjoint run in ``libpm_hip.so`` (``pm_diff_*`` in
``include/pm_capi.h``); the pose chain rule runs in the same library on the host in FP64
(``pm_se3_pose_grad``).  There is no counterpart in the reference: the forward stages are the
reference's cost (``motion_stereo/calc_cost.cu:11-210``) and SGM (``sgm.cu:5-113``) on float
reference intensities, followed by

    depth_hat(p) = sum_d softmax_d(-agg(p, d) / (cost_scale * temperature)) / idep_d
    loss = mean over M of |depth_hat - gt|,   M = pixels with a valid cost and gt > 0.

Cost normalisation: the aggregate is in the reference's units (0..255 intensity differences,
summed over the paths); ``cost_scale`` (default 255, one full-scale intensity step) divides it
before the temperature, so T = 0.05 means logits of agg / 12.75.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi as C
from .pipeline import _hom_array, _ptr, _stream


def _d(a, n):
    a = np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1))
    assert a.size == n
    return a


def se3_homography(K, R0, t0, xi):
    """FP64 (H 3x3, t_h 3) of the pose Exp(xi) (R0, t0); xi = (rho, phi), translation first."""
    Kd, Rd, td, xd = _d(K, 9), _d(R0, 9), _d(t0, 3), _d(xi, 6)
    H = np.zeros(9)
    t = np.zeros(3)
    C.check(C.load().pm_se3_homography(C.dptr(Kd), C.dptr(Rd), C.dptr(td), C.dptr(xd),
                                       C.dptr(H), C.dptr(t)))
    return H.reshape(3, 3), t


def se3_pose_grad(K, R0, t0, xi, grad_hom):
    """dL/dxi (6) from dL/d(h11..h33, t1..t3) (12), FP64."""
    Kd, Rd, td, xd, gd = _d(K, 9), _d(R0, 9), _d(t0, 3), _d(xi, 6), _d(grad_hom, 12)
    out = np.zeros(6)
    C.check(C.load().pm_se3_pose_grad(C.dptr(Kd), C.dptr(Rd), C.dptr(td), C.dptr(xd),
                                      C.dptr(gd), C.dptr(out)))
    return out


class SoftArgminDepth:
    """One ``pm_diff`` handle: device buffers for a W x H x D keyframe with S sources.

    ``recompute=True`` keeps only the cost volume, the source counts and the readout between
    forward and backward (the aggregate and the per-path L volumes are recomputed), instead of
    storing all per-path L volumes.
    """

    def __init__(self, width, height, planes, num_sources, knobs, paths=8, temperature=0.05,
                 cost_scale=255.0, recompute=False):
        self.lib = C.load()
        self.shape = (int(height), int(width), int(planes))
        self.S = int(num_sources)
        sp = C.SgmParams(float(knobs["pi1"]), float(knobs["pi2"]), float(knobs["tau_so"]),
                         float(knobs["sgm_q1"]), float(knobs["sgm_q2"]), int(paths))
        dp = C.DiffParams(float(temperature), float(cost_scale), 1 if recompute else 0)
        h = ctypes.c_void_p()
        C.check(self.lib.pm_diff_create(C.Shape(int(width), int(height), int(planes)), self.S,
                                        sp, dp, ctypes.byref(h)))
        self._h = h
        self._keep = None

    @property
    def memory_bytes(self) -> int:
        b = ctypes.c_uint64()
        C.check(self.lib.pm_diff_memory(self._h, ctypes.byref(b)))
        return int(b.value)

    def forward(self, ref, quads, homs, idep_min, dep_sample, gt):
        """ref: float32 [H, W] device; quads: S quad textures; homs: S (h, t) float32 pairs;
        gt: float32 [H, W] device.  Returns (depth_hat [H, W] float32, loss [1] float64), both
        on the device."""
        H, W, _ = self.shape
        assert ref.dtype == torch.float32 and tuple(ref.shape) == (H, W)
        assert gt.dtype == torch.float32 and tuple(gt.shape) == (H, W)
        assert len(quads) == self.S and len(homs) == self.S
        ref = ref.contiguous()
        gt = gt.contiguous()
        depth = torch.empty((H, W), dtype=torch.float32, device=ref.device)
        loss = torch.empty((1,), dtype=torch.float64, device=ref.device)
        qarr = (ctypes.c_void_p * self.S)(*[q.data_ptr() for q in quads])
        C.check(self.lib.pm_diff_forward(self._h, _ptr(ref), qarr, _hom_array(homs),
                                         float(idep_min), float(dep_sample), _ptr(gt),
                                         _ptr(depth), _ptr(loss), _stream()))
        self._keep = (ref, list(quads), gt, depth)   # read again by the backward
        return depth, loss

    def backward(self, grad_loss=1.0):
        """Returns (dL/dref [H, W] float32, dL/d(H, t) [S, 12] float64), on the device.  In the
        store mode it consumes the forward's stored volumes: another backward needs another
        forward."""
        H, W, _ = self.shape
        dev = self._keep[0].device
        gref = torch.empty((H, W), dtype=torch.float32, device=dev)
        ghom = torch.empty((self.S, 12), dtype=torch.float64, device=dev)
        C.check(self.lib.pm_diff_backward(self._h, float(grad_loss), _ptr(gref), _ptr(ghom),
                                          _stream()))
        return gref, ghom

    def download(self, name):
        """Host copy of a handle buffer: "cost", "count", "agg" (dL/dagg after backward),
        "grad_cost", "valid"."""
        H, W, D = self.shape
        torch.cuda.synchronize()
        if name in ("count",):
            out = np.empty((H, W, D), np.uint8)
        elif name == "valid":
            out = np.empty((H, W), np.uint8)
        else:
            out = np.empty((H, W, D), np.float32)
        C.check(self.lib.pm_diff_download(self._h, name.encode(), out.ctypes.data_as(ctypes.c_void_p),
                                          out.nbytes))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.pm_diff_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DepthLoss(torch.autograd.Function):
    """loss = DepthLoss.apply(xi, ref, op, quads, K, poses, idep_min, dep_sample, gt)

    xi: float64 [S, 6] (CPU) se3 perturbations, (R_s, t_s) = Exp(xi_s) (R0_s, t0_s);
    ref: float32 [H, W] device reference intensities; op: a SoftArgminDepth handle (one
    forward / backward pair at a time); poses: S (R0, t0) relative poses (reference camera to
    measurement camera).  Returns the loss as a float64 scalar (CPU); the backward gives
    dL/dxi (float64, CPU) and dL/dref (float32, device).
    """

    @staticmethod
    def forward(ctx, xi, ref, op, quads, K, poses, idep_min, dep_sample, gt):
        xin = xi.detach().cpu().numpy().astype(np.float64)
        homs = []
        for s, (R0, t0) in enumerate(poses):
            Hm, th = se3_homography(K, R0, t0, xin[s])
            homs.append((Hm.astype(np.float32), th.astype(np.float32)))
        depth, loss = op.forward(ref.detach().float(), quads, homs, idep_min, dep_sample, gt)
        ctx.op, ctx.K, ctx.poses, ctx.xin = op, K, poses, xin
        ctx.ref_dtype = ref.dtype
        ctx.depth = depth
        return loss.cpu().reshape(())

    @staticmethod
    def backward(ctx, g):
        gref, ghom = ctx.op.backward(float(g))
        gh = ghom.cpu().numpy()
        gxi = np.stack([se3_pose_grad(ctx.K, R0, t0, ctx.xin[s], gh[s])
                        for s, (R0, t0) in enumerate(ctx.poses)])
        return (torch.from_numpy(gxi), gref.to(ctx.ref_dtype), None, None, None, None, None,
                None, None)
