"""Device-resident stage calls and the keyframe sequence driver.

Every function here enqueues HIP kernels from ``libpm_hip.so`` through the C ABI on torch's
current stream; torch is used only to own device memory.  Layouts are those of
``include/pm_capi.h`` (cost / aggregate pixel-major ``[H][W][D]``).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi as C

COST_TYPES = {
    "f32": (C.PM_COST_F32, torch.float32),
    "f16": (C.PM_COST_F16, torch.float16),
    "bf16": (C.PM_COST_BF16, torch.bfloat16),
}

STATE_KEYS = ("alpha", "beta", "mu", "sig")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor):
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
    return ctypes.c_void_p(t.data_ptr())


def homography(K, R_l, T_l, R_r, T_r):
    """FP64 plane-sweep homography rounded to float (stereo_mapper.cpp:79-90).

    Returns ``(h, t)``: float32 arrays of 9 (row-major) and 3 entries.
    """
    lib = C.load()
    arrs = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in
            (K, R_l, T_l, R_r, T_r)]
    out = C.Homography()
    C.check(lib.pm_homography_setup(*[C.dptr(a) for a in arrs], ctypes.byref(out)))
    return np.array(out.h[:], np.float32), np.array(out.t[:], np.float32)


def _hom_array(homs):
    arr = (C.Homography * max(len(homs), 1))()
    for i, (h, t) in enumerate(homs):
        arr[i].h[:] = [float(v) for v in np.asarray(h, np.float32).reshape(9)]
        arr[i].t[:] = [float(v) for v in np.asarray(t, np.float32).reshape(3)]
    return arr


def make_quad(img: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """uint8 [H][W] image -> bilinear quad texture (int64 [H][W], four FP16 per texel)."""
    assert img.dtype == torch.uint8 and img.dim() == 2
    H, W = img.shape
    if out is None:
        out = torch.empty((H, W), dtype=torch.int64, device=img.device)
    C.check(C.load().pm_make_quad(_ptr(img), _ptr(out), W, H, _stream()))
    return out


def calc_cost(ref: torch.Tensor, quads, homs, idep_min: float, dep_sample: float, planes: int,
              cost_type: str = "f32", out: torch.Tensor | None = None) -> torch.Tensor:
    """Plane-sweep cost over all sources -> ``[H][W][D]`` of the chosen storage type."""
    H, W = ref.shape
    code, dt = COST_TYPES[cost_type]
    if out is None:
        out = torch.empty((H, W, planes), dtype=dt, device=ref.device)
    assert out.dtype == dt and tuple(out.shape) == (H, W, planes)
    S = len(quads)
    qarr = (ctypes.c_void_p * max(S, 1))(*[q.data_ptr() for q in quads])
    C.check(C.load().pm_calc_cost(C.Shape(W, H, planes), _ptr(ref), S, qarr, _hom_array(homs),
                                  float(idep_min), float(dep_sample), _ptr(out), code, _stream()))
    return out


def _cost_code(cost: torch.Tensor) -> int:
    for code, dt in COST_TYPES.values():
        if cost.dtype == dt:
            return code
    raise TypeError(f"unsupported cost dtype {cost.dtype}")


def sgm(ref: torch.Tensor, cost: torch.Tensor, knobs: dict, paths: int,
        out: torch.Tensor | None = None) -> torch.Tensor:
    """Multi-path SGM aggregation -> float32 ``[H][W][D]``."""
    H, W, D = cost.shape
    if out is None:
        out = torch.empty((H, W, D), dtype=torch.float32, device=cost.device)
    p = C.SgmParams(knobs["pi1"], knobs["pi2"], knobs["tau_so"], knobs["sgm_q1"],
                    knobs["sgm_q2"], paths)
    C.check(C.load().pm_sgm(C.Shape(W, H, D), _ptr(ref), _ptr(cost), _cost_code(cost), p,
                            _ptr(out), _stream()))
    return out


def sgm_filter(ref: torch.Tensor, cost: torch.Tensor, knobs: dict, paths: int, idep_min: float,
               dep_sample: float, agg: torch.Tensor | None = None,
               depth: torch.Tensor | None = None):
    """SGM (paths one after another) then the WTA/subpixel filter -> (agg, depth)."""
    H, W, D = cost.shape
    if agg is None:
        agg = torch.empty((H, W, D), dtype=torch.float32, device=cost.device)
    if depth is None:
        depth = torch.empty((H, W), dtype=torch.float32, device=cost.device)
    p = C.SgmParams(knobs["pi1"], knobs["pi2"], knobs["tau_so"], knobs["sgm_q1"],
                    knobs["sgm_q2"], paths)
    C.check(C.load().pm_sgm_filter(C.Shape(W, H, D), _ptr(ref), _ptr(cost), _cost_code(cost), p,
                                   float(knobs["var_scale"]), float(idep_min), float(dep_sample),
                                   _ptr(agg), _ptr(depth), _stream()))
    return agg, depth


def sgm_workspace(H: int, W: int, D: int, paths: int, device="cuda") -> torch.Tensor:
    """Workspace of ``sgm_filter_paths``: one float32 L volume per path."""
    b = ctypes.c_uint64()
    C.check(C.load().pm_sgm_workspace_size(C.Shape(W, H, D), paths, ctypes.byref(b)))
    return torch.empty((b.value // 4,), dtype=torch.float32, device=device)


def sgm_filter_paths(ref: torch.Tensor, cost: torch.Tensor, knobs: dict, paths: int,
                     idep_min: float, dep_sample: float, workspace: torch.Tensor | None = None,
                     agg: torch.Tensor | None = None, depth: torch.Tensor | None = None,
                     write_agg: bool = True):
    """SGM with every path in flight at once (per-path volumes in ``workspace``), summed in
    path order and filtered -> (agg, depth); bit-identical to ``sgm_filter``."""
    H, W, D = cost.shape
    if workspace is None:
        workspace = sgm_workspace(H, W, D, paths, cost.device)
    if agg is None and write_agg:
        agg = torch.empty((H, W, D), dtype=torch.float32, device=cost.device)
    if depth is None:
        depth = torch.empty((H, W), dtype=torch.float32, device=cost.device)
    p = C.SgmParams(knobs["pi1"], knobs["pi2"], knobs["tau_so"], knobs["sgm_q1"],
                    knobs["sgm_q2"], paths)
    C.check(C.load().pm_sgm_filter_paths(
        C.Shape(W, H, D), _ptr(ref), _ptr(cost), _cost_code(cost), p, float(knobs["var_scale"]),
        float(idep_min), float(dep_sample), _ptr(workspace), workspace.numel() * 4,
        _ptr(agg) if agg is not None else None, _ptr(depth), _stream()))
    return agg, depth


def sgm_excess_workspace(H: int, W: int, D: int, paths: int, device="cuda") -> torch.Tensor:
    """Workspace of ``sgm_filter_excess``: one uint16 excess volume per path."""
    b = ctypes.c_uint64()
    C.check(C.load().pm_sgm_excess_workspace_size(C.Shape(W, H, D), paths, ctypes.byref(b)))
    return torch.empty((b.value // 2,), dtype=torch.int16, device=device)


def sgm_filter_excess(ref: torch.Tensor, cost: torch.Tensor, knobs: dict, paths: int,
                      idep_min: float, dep_sample: float, workspace: torch.Tensor | None = None,
                      agg: torch.Tensor | None = None, depth: torch.Tensor | None = None,
                      write_agg: bool = True):
    """Compact SGM: every path in flight at once, each storing its 16-bit excess over the
    cost; the aggregate is rebuilt from the cost and the excess volumes and filtered ->
    (agg, depth)."""
    H, W, D = cost.shape
    if workspace is None:
        workspace = sgm_excess_workspace(H, W, D, paths, cost.device)
    if agg is None and write_agg:
        agg = torch.empty((H, W, D), dtype=torch.float32, device=cost.device)
    if depth is None:
        depth = torch.empty((H, W), dtype=torch.float32, device=cost.device)
    p = C.SgmParams(knobs["pi1"], knobs["pi2"], knobs["tau_so"], knobs["sgm_q1"],
                    knobs["sgm_q2"], paths)
    C.check(C.load().pm_sgm_filter_excess(
        C.Shape(W, H, D), _ptr(ref), _ptr(cost), _cost_code(cost), p, float(knobs["var_scale"]),
        float(idep_min), float(dep_sample), _ptr(workspace), workspace.numel() * 2,
        _ptr(agg) if agg is not None else None, _ptr(depth), _stream()))
    return agg, depth


def filter_cost(agg: torch.Tensor, var_scale: float, idep_min: float, dep_sample: float,
                out: torch.Tensor | None = None) -> torch.Tensor:
    H, W, D = agg.shape
    if out is None:
        out = torch.empty((H, W), dtype=torch.float32, device=agg.device)
    C.check(C.load().pm_filter_cost(C.Shape(W, H, D), _ptr(agg), float(var_scale),
                                    float(idep_min), float(dep_sample), _ptr(out), _stream()))
    return out


def wta(cost: torch.Tensor, idep_min: float, dep_sample: float) -> torch.Tensor:
    H, W, D = cost.shape
    out = torch.empty((H, W), dtype=torch.float32, device=cost.device)
    C.check(C.load().pm_wta(C.Shape(W, H, D), _ptr(cost), _cost_code(cost), float(idep_min),
                            float(dep_sample), _ptr(out), _stream()))
    return out


def median3x3(depth: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    H, W = depth.shape
    if out is None:
        out = torch.empty_like(depth)
    C.check(C.load().pm_median3x3(_ptr(depth), _ptr(out), W, H, _stream()))
    return out


def new_state(H: int, W: int, device) -> dict:
    return {k: torch.zeros((H, W), dtype=torch.float32, device=device) for k in STATE_KEYS}


def _state(s: dict) -> C.FusionState:
    return C.FusionState(*[s[k].data_ptr() for k in STATE_KEYS])


def depth_fuse(depth: torch.Tensor, zbuf: torch.Tensor, pre: dict, cur: dict, r_p2c, t_p2c,
               cam, table: torch.Tensor | None = None, fused: torch.Tensor | None = None,
               zbuf_next: torch.Tensor | None = None) -> None:
    """One Gaussian x Beta fusion step on ``pre`` and ``cur``; the fused depth goes to
    ``fused`` (default: ``depth`` itself, in place).  Without ``zbuf_next`` the step empties
    ``zbuf`` first; with it, ``zbuf`` must be empty and ``zbuf_next`` is emptied for the next
    step (see :func:`zbuf_reset`)."""
    H, W = depth.shape
    r = (ctypes.c_float * 9)(*[float(v) for v in np.asarray(r_p2c, np.float32).reshape(9)])
    t = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(t_p2c, np.float32).reshape(3)])
    tab = _ptr(table) if table is not None else ctypes.c_void_p(0)
    out = fused if fused is not None else depth
    nxt = _ptr(zbuf_next) if zbuf_next is not None else ctypes.c_void_p(0)
    C.check(C.load().pm_depth_fuse(W, H, C.Camera(*[float(c) for c in cam]), r, t, _ptr(depth),
                                   _ptr(out), _ptr(zbuf), nxt, tab, _state(pre), _state(cur),
                                   _stream()))


def zbuf_reset(zbuf: torch.Tensor) -> None:
    """Empties a fusion z-buffer (int64 tensor) on the current stream."""
    C.check(C.load().pm_zbuf_reset(_ptr(zbuf), zbuf.numel(), _stream()))


def depth_fuse_cpu_rules(depth: torch.Tensor, pre: dict, cur: dict, r_p2c, t_p2c, cam,
                         table: torch.Tensor | None = None,
                         work: torch.Tensor | None = None) -> torch.Tensor:
    """outputFusionCPU's CASE >= 3 step (the reference's CPU rules), in place on ``depth``
    (filtered depth in, result out), ``pre`` and ``cur``; r_p2c, t_p2c in double.  Returns the
    scratch it used (pass it back to reuse it)."""
    H, W = depth.shape
    need = ctypes.c_uint64(0)
    C.check(C.load().pm_depth_fuse_cpu_rules_workspace_size(W, H, ctypes.byref(need)))
    if work is None or work.numel() < need.value:
        work = torch.empty(need.value, dtype=torch.uint8, device=depth.device)
    r = np.ascontiguousarray(np.asarray(r_p2c, np.float64).reshape(9))
    t = np.ascontiguousarray(np.asarray(t_p2c, np.float64).reshape(3))
    tab = _ptr(table) if table is not None else ctypes.c_void_p(0)
    C.check(C.load().pm_depth_fuse_cpu_rules(
        W, H, C.Camera(*[float(c) for c in cam]), C.dptr(r), C.dptr(t), _ptr(depth), tab,
        _ptr(work), work.numel(), _state(pre), _state(cur), _stream()))
    return work


def pose_p2c(R, t, R_prev, t_prev):
    """R_p2c = R^T R_prev, t_p2c = R^T (t_prev - t) in FP64, cast to float
    (stereo_mapper.cpp:155-156, 167-170)."""
    R = np.asarray(R, np.float64); R_prev = np.asarray(R_prev, np.float64)
    t = np.asarray(t, np.float64).reshape(3); t_prev = np.asarray(t_prev, np.float64).reshape(3)
    return (R.T @ R_prev).astype(np.float32), (R.T @ (t_prev - t)).astype(np.float32)


class KeyframeSequence:
    """Device-resident driver of a keyframe sequence (the StereoMapper::outputFusion loop).

    All frames are uploaded once and turned into quad textures; each keyframe then runs
    cost (over its sources) -> SGM -> filter (-> median) -> fusion on the current stream with
    no host synchronisation.  The first keyframe only resets the filter state
    (stereo_mapper.cpp:128-146); later ones fuse into the ping-pong state (:149-243).
    ``lookahead``: each fusion step also runs the next keyframe's forward warp (see
    ``fusion_stage``).
    """

    def __init__(self, images, poses, sources, K, planes, idep_min, idep_max, knobs, paths,
                 cost_type="f32", median=False, device="cuda", sgm_storage="exact",
                 write_agg=True, lookahead=False):
        self.device = torch.device(device)
        self.lookahead = bool(lookahead)
        self.H, self.W = np.asarray(images[0]).shape
        self.D = planes
        self.idep_min = float(idep_min)
        self.dep_sample = float((idep_max - idep_min) / (planes - 1))
        self.knobs = dict(knobs)
        self.paths = paths
        self.cost_type = cost_type
        self.median = median
        # "sweep" exists only in the batched form (BatchedSequence)
        assert sgm_storage in ("exact", "compact", "sweep"), sgm_storage
        self.sgm_storage = sgm_storage
        self.write_agg = write_agg   # materialise the aggregate (the filter needs only depth)
        self.K = np.asarray(K, np.float64)
        self.cam = (self.K[0, 0], self.K[1, 1], self.K[0, 2], self.K[1, 2])
        self.poses = [(np.asarray(R, np.float64), np.asarray(T, np.float64).reshape(3))
                      for R, T in poses]
        self.sources = [list(s) for s in sources]
        self.images = [torch.from_numpy(np.ascontiguousarray(im)).to(self.device)
                       for im in images]
        self.quads = [make_quad(im) for im in self.images]
        self.homs = []
        for i, srcs in enumerate(self.sources):
            Rl, Tl = self.poses[i]
            self.homs.append([homography(self.K, Rl, Tl, *self.poses[j]) for j in srcs])
        H, W, D = self.H, self.W, self.D
        dt = COST_TYPES[cost_type][1]
        self.cost = torch.empty((H, W, D), dtype=dt, device=self.device)
        self.agg = (torch.empty((H, W, D), dtype=torch.float32, device=self.device)
                    if write_agg else None)
        if sgm_storage == "compact":
            self.sgm_ws = sgm_excess_workspace(H, W, D, paths, self.device)
        elif sgm_storage == "exact":
            self.sgm_ws = sgm_workspace(H, W, D, paths, self.device)
        else:
            self.sgm_ws = None
        self.dep = torch.empty((H, W), dtype=torch.float32, device=self.device)
        self.dep_raw = torch.empty((H, W), dtype=torch.float32, device=self.device)
        self.fused = torch.empty((H, W), dtype=torch.float32, device=self.device)
        # three z-buffers in rotation: fusion step k reads zbufs[k % 3], fills (or lets the
        # next step's warp fill) zbufs[(k + 1) % 3] and empties the other one for later
        self.zbufs = [torch.empty((H, W), dtype=torch.int64, device=self.device)
                      for _ in range(3)]
        self.table = torch.empty((H, W), dtype=torch.int32, device=self.device)
        self.states = [new_state(H, W, self.device), new_state(H, W, self.device)]
        self.reset()

    def reset(self):
        self.first = True
        self.pre = 0
        self.prev_pose = None
        self.prev_index = None
        self.zk = 0            # fusion steps so far (z-buffer rotation)
        self.warped = None     # keyframe whose warp the previous step already ran (lookahead)
        for z in self.zbufs:
            zbuf_reset(z)

    def _cost_args(self, i: int):
        """ctypes arguments of keyframe i's pm_calc_cost (built once; the stream is added per
        call)."""
        if not hasattr(self, "_cost_cache"):
            self._cost_cache = {}
        if i not in self._cost_cache or self._cost_cache[i][-1] is not self.cost:
            srcs = self.sources[i]
            qarr = (ctypes.c_void_p * max(len(srcs), 1))(*[self.quads[j].data_ptr() for j in srcs])
            code = COST_TYPES[self.cost_type][0]
            self._cost_cache[i] = (C.Shape(self.W, self.H, self.D), _ptr(self.images[i]),
                                   len(srcs), qarr, _hom_array(self.homs[i]),
                                   ctypes.c_float(self.idep_min), ctypes.c_float(self.dep_sample),
                                   _ptr(self.cost), code, self.cost)
        return self._cost_cache[i][:-1]

    def cost_stage(self, i: int) -> None:
        """Plane-sweep cost of keyframe i over its sources into ``self.cost``."""
        C.check(C.load().pm_calc_cost(*self._cost_args(i), _stream()))

    def depth_stage(self, i: int, median: bool = True) -> torch.Tensor:
        """cost -> SGM + filter for keyframe i into ``dep_raw`` (with the median) or ``dep``;
        ``median``: also run the median into ``dep`` here (``run_keyframe`` leaves it to the
        fusion step, which computes it in its own launch)."""
        self.cost_stage(i)
        out = self.dep_raw if self.median else self.dep
        self.sgm_stage(i, out)
        if self.median and median:
            median3x3(self.dep_raw, out=self.dep)
        return self.dep

    def sgm_stage(self, i: int, out: torch.Tensor) -> None:
        """SGM (every path at once) + filter of keyframe i's cost volume into ``out``."""
        if self.sgm_storage == "sweep":
            raise ValueError("the sweep SGM storage runs in the batched form only")
        if not hasattr(self, "_sgm_cache"):
            self._sgm_cache = {}
        key = (i, out.data_ptr(), self.cost.data_ptr())
        if key not in self._sgm_cache:
            k = self.knobs
            p = C.SgmParams(k["pi1"], k["pi2"], k["tau_so"], k["sgm_q1"], k["sgm_q2"], self.paths)
            elt = 2 if self.sgm_storage == "compact" else 4
            self._sgm_cache[key] = (
                C.Shape(self.W, self.H, self.D), _ptr(self.images[i]), _ptr(self.cost),
                _cost_code(self.cost), p, ctypes.c_float(k["var_scale"]),
                ctypes.c_float(self.idep_min), ctypes.c_float(self.dep_sample),
                _ptr(self.sgm_ws), ctypes.c_uint64(self.sgm_ws.numel() * elt),
                _ptr(self.agg) if self.agg is not None else None, _ptr(out))
        fn = (C.load().pm_sgm_filter_excess if self.sgm_storage == "compact"
              else C.load().pm_sgm_filter_paths)
        C.check(fn(*self._sgm_cache[key], _stream()))

    def _measurement(self, i: int) -> torch.Tensor:
        """Keyframe i's filtered depth as the filter wrote it (before the median, if any)."""
        return self.dep_raw if self.median else self.dep

    def _median_out(self) -> torch.Tensor:
        return self.dep

    def fusion_stage(self, i: int) -> None:
        """(The median, then) one outputFusion step of keyframe i (stereo_mapper.cpp:128-243):
        the forward warp of the previous state (pm_depth_fuse_warp) and the update with the
        median in the same launch (pm_depth_fuse_step).  With ``lookahead`` the update also
        runs keyframe i+1's warp (its pose is known), so each later step is one launch; the
        state's mu / sig then hold the values warped into keyframe i+1 until that step runs
        (the fused depth, the table and the other state fields are those of keyframe i)."""
        R, t = self.poses[i]
        raw = self._measurement(i)
        out = self._median_out() if self.median else raw
        if self.first:
            self.first = False
            self.pre = 0
            for s in self.states:
                s["alpha"].zero_()
            if self.median:
                median3x3(raw, out=out)
            self.dep = out
            self.prev_pose = (R, t)
            self.prev_index = i
            self.fused.zero_()
            return
        if self.warped is not None and self.warped != i:
            raise ValueError(f"fusion_stage({i}): the previous step warped keyframe "
                             f"{self.warped}")
        nxt = self.lookahead and i + 1 < len(self.images)
        k = self.zk % 3
        key = (i, self.prev_index, self.pre, k, raw.data_ptr(), out.data_ptr(), nxt,
               self.warped is None)
        if not hasattr(self, "_fuse_cache"):
            self._fuse_cache = {}
        if key not in self._fuse_cache:
            def f32(a, n):
                return (ctypes.c_float * n)(*[float(v) for v in
                                              np.asarray(a, np.float32).reshape(n)])
            cam = C.Camera(*[float(c) for c in self.cam])
            r, tt = pose_p2c(R, t, *self.prev_pose)
            pre, cur = self.states[self.pre], self.states[self.pre ^ 1]
            z = [_ptr(self.zbufs[(k + d) % 3]) for d in range(3)]
            warp = None if self.warped is not None else \
                (self.W, self.H, cam, f32(r, 9), f32(tt, 3), _state(pre), z[0], z[1])
            rn = tn = None
            if nxt:
                r2, t2 = pose_p2c(*self.poses[i + 1], R, t)
                rn, tn = f32(r2, 9), f32(t2, 3)
            step = (self.W, self.H, cam, _ptr(raw), int(self.median),
                    _ptr(out) if self.median else None, _ptr(self.fused), z[0],
                    _ptr(self.table), _state(pre), _state(cur), rn, tn,
                    z[1] if nxt else None, z[2] if nxt else None)
            self._fuse_cache[key] = (warp, step)
        warp, step = self._fuse_cache[key]
        lib = C.load()
        if warp is not None:
            C.check(lib.pm_depth_fuse_warp(*warp, _stream()))
        C.check(lib.pm_depth_fuse_step(*step, _stream()))
        self.dep = out
        self.warped = i + 1 if nxt else None
        self.prev_pose = (R, t)
        self.prev_index = i
        self.pre ^= 1
        self.zk += 1

    def run_keyframe(self, i: int) -> None:
        self.depth_stage(i, median=False)
        self.fusion_stage(i)

    def run(self) -> None:
        self.reset()
        for i in range(len(self.images)):
            self.run_keyframe(i)

    @property
    def state(self) -> dict:
        """The state written by the last fusion step."""
        return self.states[self.pre]


def keyframe_spec(kf: dict, device="cuda") -> dict:
    """A ``Batch`` keyframe from host data: ``ref`` and ``srcs`` (uint8 images), ``K``,
    ``ref_pose`` and ``src_poses`` (camera->world R, T), ``planes``, ``idep_min`` and
    ``idep_max``.  Uploads the frames, builds the sources' quad textures and the FP64
    homography setup of each source (stereo_mapper.cpp:79-90)."""
    ref = torch.from_numpy(np.ascontiguousarray(kf["ref"])).to(device)
    quads = [make_quad(torch.from_numpy(np.ascontiguousarray(im)).to(device))
             for im in kf["srcs"]]
    homs = [homography(kf["K"], *kf["ref_pose"], R, T) for R, T in kf["src_poses"]]
    D = int(kf["planes"])
    return {"ref": ref, "quads": quads, "homs": homs, "planes": D,
            "idep_min": float(kf["idep_min"]),
            "dep_sample": float((kf["idep_max"] - kf["idep_min"]) / (D - 1))}


class Batch:
    """A ``pm_batch``: cost -> SGM -> filter for many keyframes of uneven size per launch.

    ``keyframes`` is a list of dicts with ``ref`` (uint8 device tensor), ``quads`` (device quad
    textures of the measurement frames), ``homs`` ([(h, t)]), ``planes``, ``idep_min``,
    ``dep_sample``.  Output tensors ``cost``, ``agg`` and ``depth`` are allocated per keyframe.
    """

    def __init__(self, keyframes, cost_type, knobs, paths, storage="exact", write_agg=True):
        self.lib = C.load()
        self.cost_type = cost_type
        self.storage = storage
        st = {"exact": C.PM_SGM_EXACT, "compact": C.PM_SGM_COMPACT,
              "sweep": C.PM_SGM_SWEEP}[storage]
        code, dt = COST_TYPES[cost_type]
        n = len(keyframes)
        self.kfs = keyframes
        arr = (C.Keyframe * n)()
        self.outputs = []
        for i, kf in enumerate(keyframes):
            H, W = kf["ref"].shape
            D = kf["planes"]
            dev = kf["ref"].device
            # the exact storage accumulates into agg; compact may skip writing it
            keep = write_agg or storage == "exact"
            out = {"cost": torch.empty((H, W, D), dtype=dt, device=dev),
                   "agg": (torch.empty((H, W, D), dtype=torch.float32, device=dev)
                           if keep else None),
                   "depth": torch.empty((H, W), dtype=torch.float32, device=dev)}
            self.outputs.append(out)
            k = arr[i]
            k.width, k.height, k.planes = W, H, D
            k.num_sources = len(kf["quads"])
            k.idep_min, k.dep_sample = float(kf["idep_min"]), float(kf["dep_sample"])
            k.ref = kf["ref"].data_ptr()
            for s, (q, (h, t)) in enumerate(zip(kf["quads"], kf["homs"])):
                k.src_quads[s] = q.data_ptr()
                k.homs[s].h[:] = [float(v) for v in np.asarray(h, np.float32).reshape(9)]
                k.homs[s].t[:] = [float(v) for v in np.asarray(t, np.float32).reshape(3)]
            k.cost = out["cost"].data_ptr()
            k.agg = out["agg"].data_ptr() if out["agg"] is not None else None
            k.depth = out["depth"].data_ptr()
        p = C.SgmParams(knobs["pi1"], knobs["pi2"], knobs["tau_so"], knobs["sgm_q1"],
                        knobs["sgm_q2"], paths)
        h = ctypes.c_void_p()
        C.check(self.lib.pm_batch_create(arr, n, code, p, float(knobs["var_scale"]), st,
                                         ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self.lib.pm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def tuning(self) -> dict:
        t = C.BatchTuning()
        C.check(self.lib.pm_batch_get_tuning(self._h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in C.BatchTuning._fields_}

    @tuning.setter
    def tuning(self, values: dict):
        t = C.BatchTuning(**{**self.tuning, **values})
        C.check(self.lib.pm_batch_set_tuning(self._h, ctypes.byref(t)))

    def run(self):
        C.check(self.lib.pm_batch_run(self._h, _stream()))

    def run_stage(self, stage: int):
        """0 = cost, 1 = SGM (all paths), 2 = filter."""
        C.check(self.lib.pm_batch_run_stage(self._h, int(stage), _stream()))

    def autotune(self, reps: int = 3) -> dict:
        t = C.BatchTuning()
        C.check(self.lib.pm_batch_autotune(self._h, _stream(), reps, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in C.BatchTuning._fields_}

    def capture(self):
        C.check(self.lib.pm_batch_capture(self._h, _stream()))

    def launch(self):
        C.check(self.lib.pm_batch_launch(self._h, _stream()))


# Keyframes per batch group: the cost volumes of one group stay within this many bytes, so
# that the group's SGM pass re-reads them from the 256 MB Infinity Cache (measured on the
# 752x480x64 FP16 sequence: groups of 5 keyframes, 230 MB, 2.8 % faster than one batch of 20;
# groups of 2 slower).
GROUP_CACHE_BYTES = 240 << 20


def cache_groups(cost_bytes, limit: int = GROUP_CACHE_BYTES, min_group: int = 4) -> list[range]:
    """Consecutive keyframe ranges whose cost volumes sum to at most ``limit`` bytes, but with
    at least ``min_group`` keyframes per group (smaller launches lose more to their tails)."""
    groups, start, acc = [], 0, 0
    for i, b in enumerate(cost_bytes):
        if i - start >= min_group and acc + b > limit:
            groups.append(range(start, i))
            start, acc = i, 0
        acc += b
    if start < len(cost_bytes):
        groups.append(range(start, len(cost_bytes)))
    return groups


class BatchGroups:
    """Several ``Batch`` objects run one after another with one interface: ``outputs`` lists
    every keyframe in order, and tuning, capture, launch and the stage calls apply to all."""

    def __init__(self, batches):
        self.batches = batches
        self.outputs = [o for b in batches for o in b.outputs]

    def close(self):
        for b in self.batches:
            b.close()

    @property
    def tuning(self) -> dict:
        """The first group's launch shapes (``tunings`` lists every group's)."""
        return self.batches[0].tuning

    @property
    def tunings(self) -> list[dict]:
        return [b.tuning for b in self.batches]

    @tuning.setter
    def tuning(self, values: dict):
        for b in self.batches:
            b.tuning = values

    def autotune(self, reps: int = 3) -> dict:
        """Tunes every group on its own keyframes (groups differ in size, source and plane
        counts, and a shape valid for one may be refused for another); returns the first
        group's choice."""
        best = [b.autotune(reps) for b in self.batches]
        return best[0]

    def run(self):
        for b in self.batches:
            b.run()

    def run_stage(self, stage: int):
        for b in self.batches:
            b.run_stage(stage)

    def capture(self):
        for b in self.batches:
            b.capture()

    def launch(self):
        for b in self.batches:
            b.launch()


class BatchedSequence(KeyframeSequence):
    """The keyframe sequence with the keyframes' cost, SGM and filter batched, then the
    fusion steps in order (fusion is sequential in time, stereo_mapper.cpp:128-243).

    ``group``: keyframes per batch; "cache" (default) sizes groups by ``cache_groups`` so each
    group's cost volumes stay cache-resident for its SGM pass, an integer fixes the size, and
    0 puts every keyframe in one batch.  The sweep storage always takes one batch: each of its
    launches covers one 16-row block of every keyframe, so it needs every keyframe in the
    launch to fill the chip (groups of 5 at 752x480 give 60 workgroups for 256 CUs)."""

    def __init__(self, *args, group="cache", **kw):
        super().__init__(*args, **kw)
        H, W, D = self.H, self.W, self.D
        self.cost = self.agg = self.sgm_ws = None   # per keyframe buffers live in the batch
        torch.cuda.empty_cache()
        n = len(self.images)
        specs = [{"ref": self.images[i], "quads": [self.quads[j] for j in self.sources[i]],
                  "homs": self.homs[i], "planes": D, "idep_min": self.idep_min,
                  "dep_sample": self.dep_sample} for i in range(n)]
        if self.sgm_storage == "sweep":
            group = 0
        if group == "cache":
            per = H * W * D * COST_TYPES[self.cost_type][1].itemsize
            ranges = cache_groups([per] * n) or [range(0, n)]
        elif not group:
            ranges = [range(0, n)]
        else:
            ranges = [range(k, min(n, k + int(group))) for k in range(0, n, int(group))]
        batches = [Batch([specs[i] for i in r], self.cost_type, self.knobs, self.paths,
                         self.sgm_storage, write_agg=self.write_agg) for r in ranges]
        self.batch = batches[0] if len(batches) == 1 else BatchGroups(batches)
        self.dep_med = torch.empty((H, W), dtype=torch.float32, device=self.device) \
            if self.median else None

    def depth_of(self, i: int) -> torch.Tensor:
        """Keyframe i's filtered depth as the batch wrote it (before the median, if any)."""
        return self.batch.outputs[i]["depth"]

    def _measurement(self, i: int) -> torch.Tensor:
        return self.depth_of(i)

    def _median_out(self) -> torch.Tensor:
        return self.dep_med

    def fusion_step(self, i: int) -> None:
        """(median ->) fusion of keyframe i, after the batch has run."""
        self.fusion_stage(i)

    def run(self) -> None:
        self.reset()
        self.batch.run()
        for i in range(len(self.images)):
            self.fusion_step(i)

    def run_overlapped(self) -> None:
        """``run`` with group g+1's cost launch beside group g's SGM pass: the current stream
        runs the groups' cost launches one after another, and a second stream runs each group's
        SGM pass, filter and fusion steps once that group's cost is written (the cost kernel is
        VALU-bound, the SGM pass bound by memory).  The current stream waits for the second one
        at the end, so every result of the sequence is ready there; the results equal
        ``run``'s."""
        main = torch.cuda.current_stream()
        if getattr(self, "_side", None) is None:
            self._side = torch.cuda.Stream(device=self.device)
        side = self._side
        groups = getattr(self.batch, "batches", [self.batch])
        self.reset()
        side.wait_stream(main)
        k = 0
        for b in groups:
            b.run_stage(0)
            done = torch.cuda.Event()
            done.record(main)
            with torch.cuda.stream(side):
                side.wait_event(done)
                b.run_stage(1)
                b.run_stage(2)
                for i in range(k, k + len(b.outputs)):
                    self.fusion_step(i)
            k += len(b.outputs)
        main.wait_stream(side)

    def capture_overlapped(self) -> None:
        """Records ``run_overlapped`` (both streams) into one hipGraph, after one eager run."""
        cap = torch.cuda.Stream(device=self.device)
        cap.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(cap):
            self.run_overlapped()
        torch.cuda.current_stream().wait_stream(cap)
        self._graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._graph, stream=cap):
            self.run_overlapped()

    def launch_overlapped(self) -> None:
        """Replays the graph of ``capture_overlapped`` on the current stream."""
        self._graph.replay()
