"""C-ABI checks that need no GPU: every declared symbol is exported, argument validation
returns status codes (never exits), and the FP64 homography setup is right."""
import ctypes
import os
import re

import numpy as np
import pytest

from probabilistic_mapping_amd import _capi as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in ("pm_capi.h",):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    lib = C.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(C.SIGNATURES), "ctypes bindings out of sync with the header"


def test_knob_globals_have_reference_defaults():
    lib = C.load()
    vals = {n: ctypes.c_float.in_dll(lib, n).value for n in
            ("pi1", "pi2", "tau_so", "sgm_q1", "sgm_q2", "var_scale")}
    assert vals == pytest.approx({"pi1": 8, "pi2": 24, "tau_so": 4, "sgm_q1": 2, "sgm_q2": 3,
                                  "var_scale": 1.02})
    assert ctypes.c_int.in_dll(lib, "CASE").value == 0


def test_knob_roundtrip():
    from probabilistic_mapping_amd.mapper import get_knobs, set_knobs
    old = set_knobs(pi1=3.5, sgm_paths=8, idep_min=0.125)
    try:
        k = get_knobs()
        assert k["pi1"] == 3.5 and k["sgm_paths"] == 8 and k["idep_min"] == 0.125
    finally:
        set_knobs(**old)
    assert get_knobs() == old


def _shape(w=8, h=6, d=4):
    return C.Shape(w, h, d)


NULL = ctypes.c_void_p(0)
FAKE = ctypes.c_void_p(0x1000)   # never dereferenced: validation fails before any launch


@pytest.mark.parametrize("call,expect", [
    (lambda L: L.pm_calc_cost(_shape(), NULL, 0, None, None, 0.0, 0.1, FAKE, 0, NULL), 1),
    (lambda L: L.pm_calc_cost(_shape(d=0), FAKE, 0, None, None, 0.0, 0.1, FAKE, 0, NULL), 1),
    (lambda L: L.pm_calc_cost(_shape(d=257), FAKE, 0, None, None, 0.0, 0.1, FAKE, 0, NULL), 1),
    (lambda L: L.pm_calc_cost(_shape(), FAKE, 17, None, None, 0.0, 0.1, FAKE, 0, NULL), 1),
    (lambda L: L.pm_calc_cost(_shape(), FAKE, 0, None, None, float("nan"), 0.1, FAKE, 0, NULL), 1),
    (lambda L: L.pm_calc_cost(_shape(), FAKE, 0, None, None, 0.0, 0.1, FAKE, 7, NULL), 1),
    (lambda L: L.pm_sgm(_shape(), FAKE, FAKE, 0, C.SgmParams(8, 24, 4, 2, 3, 5), FAKE, NULL), 1),
    (lambda L: L.pm_sgm(_shape(), FAKE, FAKE, 0, C.SgmParams(-1, 24, 4, 2, 3, 4), FAKE, NULL), 1),
    (lambda L: L.pm_sgm(_shape(w=0), FAKE, FAKE, 0, C.SgmParams(8, 24, 4, 2, 3, 4), FAKE, NULL), 1),
    (lambda L: L.pm_filter_cost(_shape(), NULL, 1.02, 0.0, 0.1, FAKE, NULL), 1),
    (lambda L: L.pm_median3x3(FAKE, FAKE, 8, 6, NULL), 1),
    (lambda L: L.pm_make_quad(FAKE, FAKE, -3, 6, NULL), 1),
    (lambda L: L.pm_wta(_shape(), FAKE, 9, 0.0, 0.1, FAKE, NULL), 1),
])
def test_argument_validation(call, expect):
    lib = C.load()
    assert call(lib) == expect
    assert C.last_error() != ""


def test_depth_fuse_validation():
    lib = C.load()
    r = (ctypes.c_float * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    t = (ctypes.c_float * 3)(0, 0, 0)
    st = C.FusionState(0x1000, 0x2000, 0x3000, 0x4000)
    cam = C.Camera(100, 100, 3, 2)
    assert lib.pm_depth_fuse(8, 6, cam, r, t, FAKE, FAKE, FAKE, NULL, NULL, st, st,
                             NULL) == 1   # state alias
    st2 = C.FusionState(0x5000, 0x6000, 0x7000, 0x8000)
    assert lib.pm_depth_fuse(8, 6, C.Camera(0, 100, 3, 2), r, t, FAKE, FAKE, FAKE, NULL, NULL,
                             st, st2, NULL) == 1
    bad = (ctypes.c_float * 9)(1, 0, 0, 0, float("inf"), 0, 0, 0, 1)
    assert lib.pm_depth_fuse(8, 6, cam, bad, t, FAKE, FAKE, FAKE, NULL, NULL, st, st2,
                             NULL) == 1
    # the next step's z-buffer must be a second buffer; a null output is refused
    assert lib.pm_depth_fuse(8, 6, cam, r, t, FAKE, FAKE, FAKE, FAKE, NULL, st, st2,
                             NULL) == 1
    assert lib.pm_depth_fuse(8, 6, cam, r, t, FAKE, NULL, FAKE, NULL, NULL, st, st2,
                             NULL) == 1
    assert lib.pm_zbuf_reset(NULL, 4, NULL) == 1


def test_depth_fuse_warp_and_step_validation():
    """pm_depth_fuse_warp / pm_depth_fuse_step refuse bad arguments before any launch."""
    lib = C.load()
    P = ctypes.c_void_p
    r = (ctypes.c_float * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    t = (ctypes.c_float * 3)(0, 0, 0)
    bad = (ctypes.c_float * 9)(1, 0, 0, 0, float("nan"), 0, 0, 0, 1)
    st = C.FusionState(0x1000, 0x2000, 0x3000, 0x4000)
    st2 = C.FusionState(0x5000, 0x6000, 0x7000, 0x8000)
    half = C.FusionState(0x1000, 0x2000, 0, 0x4000)
    cam = C.Camera(100, 100, 3, 2)
    dep, out, fused = P(0x10000), P(0x20000), P(0x30000)
    z0, z1, z2, tab = P(0x40000), P(0x50000), P(0x60000), P(0x70000)
    warp = lib.pm_depth_fuse_warp
    assert warp(8, 6, cam, r, t, st, NULL, z1, NULL) == 1                  # no z-buffer
    assert warp(8, 6, cam, r, t, st, z0, z0, NULL) == 1                    # clears its own
    assert warp(8, 6, C.Camera(100, 0, 3, 2), r, t, st, z0, z1, NULL) == 1
    assert warp(8, 6, cam, bad, t, st, z0, z1, NULL) == 1
    assert warp(8, 6, cam, r, t, half, z0, z1, NULL) == 1
    assert warp(0, 6, cam, r, t, st, z0, z1, NULL) == 1
    step = lib.pm_depth_fuse_step
    # state alias; null output; bad size
    assert step(8, 6, cam, dep, 1, out, fused, z0, tab, st, st, None, None, NULL, NULL,
                NULL) == 1
    assert step(8, 6, cam, dep, 1, out, NULL, z0, tab, st, st2, None, None, NULL, NULL,
                NULL) == 1
    assert step(8, 70000, cam, dep, 1, out, fused, z0, tab, st, st2, None, None, NULL, NULL,
                NULL) == 1
    # the median reads neighbours: its input must not alias an output
    assert step(8, 6, cam, dep, 1, out, dep, z0, tab, st, st2, None, None, NULL, NULL,
                NULL) == 1
    assert step(8, 6, cam, dep, 1, dep, fused, z0, tab, st, st2, None, None, NULL, NULL,
                NULL) == 1
    assert step(8, 6, cam, dep, 0, out, fused, z0, tab, st, st2, None, None, NULL, NULL,
                NULL) == 1                                                  # depth_out w/o median
    # the next step's warp: a complete pose, a finite one, three distinct z-buffers
    assert step(8, 6, cam, dep, 1, out, fused, z0, tab, st, st2, r, None, z1, z2, NULL) == 1
    assert step(8, 6, cam, dep, 1, out, fused, z0, tab, st, st2, r, t, NULL, z2, NULL) == 1
    assert step(8, 6, cam, dep, 1, out, fused, z0, tab, st, st2, bad, t, z1, z2, NULL) == 1
    assert step(8, 6, cam, dep, 1, out, fused, z0, tab, st, st2, r, t, z0, z2, NULL) == 1
    assert step(8, 6, cam, dep, 1, out, fused, z0, tab, st, st2, r, t, z1, z1, NULL) == 1
    assert step(8, 6, cam, dep, 1, out, fused, z0, tab, st, st2, None, None, NULL, z2,
                NULL) == 1                                                  # clear w/o next
    assert C.last_error() != ""


def test_cpu_rules_fusion_validation():
    """pm_depth_fuse_cpu_rules (outputFusionCPU CASE >= 3) refuses bad arguments before any
    launch."""
    lib = C.load()
    n = ctypes.c_uint64(0)
    assert lib.pm_depth_fuse_cpu_rules_workspace_size(8, 6, ctypes.byref(n)) == 0
    assert n.value >= 8 * 6 * 36
    assert lib.pm_depth_fuse_cpu_rules_workspace_size(0, 6, ctypes.byref(n)) == 1
    assert lib.pm_depth_fuse_cpu_rules_workspace_size(65535, 65535, ctypes.byref(n)) == 1
    r = (ctypes.c_double * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    t = (ctypes.c_double * 3)(0, 0, 0)
    cam = C.Camera(100, 100, 3, 2)
    a = C.FusionState(0x1000, 0x2000, 0x3000, 0x4000)
    b = C.FusionState(0x5000, 0x6000, 0x7000, 0x8000)
    ws = 8 * 6 * 64

    def call(w=8, h=6, cam=cam, r=r, t=t, depth=FAKE, work=FAKE, nbytes=ws, pre=a, cur=b):
        return lib.pm_depth_fuse_cpu_rules(w, h, cam, r, t, depth, NULL, work, nbytes, pre,
                                           cur, NULL)
    assert call(pre=a, cur=a) == 1                                   # pre and cur alias
    assert call(depth=NULL) == 1
    assert call(work=NULL) == 1
    assert call(nbytes=16) == 1                                      # workspace too small
    assert call(cam=C.Camera(100, 0, 3, 2)) == 1
    assert call(r=(ctypes.c_double * 9)(1, 0, 0, 0, float("nan"), 0, 0, 0, 1)) == 1
    assert call(t=(ctypes.c_double * 3)(0, float("inf"), 0)) == 1
    assert call(w=0) == 1
    assert C.last_error() != ""


def test_mapper_create_rejects_oversized_volumes():
    """A mapper whose cost volume would pass 4 GiB is refused before any device call."""
    lib = C.load()
    h = ctypes.c_void_p()
    assert lib.pm_mapper_create(65535, 65535, 256, ctypes.byref(h)) == 1 and not h.value
    assert "4 GiB" in C.last_error()
    assert lib.pm_mapper_create(8192, 8192, 16, ctypes.byref(h)) == 1 and not h.value
    assert lib.pm_mapper_create(640, 480, 0, ctypes.byref(h)) == 1 and not h.value
    assert lib.pm_mapper_create(0, 480, 64, ctypes.byref(h)) == 1 and not h.value


def test_frame_ring_and_mapper_validation():
    """The frame ring and the mapper's frame-id entry points refuse bad arguments before
    touching the device."""
    lib = C.load()
    h = ctypes.c_void_p()
    assert lib.pm_frame_ring_create(0, 6, 4, ctypes.byref(h)) == 1 and not h.value
    assert lib.pm_frame_ring_create(8, 6, 0, ctypes.byref(h)) == 1 and not h.value
    assert lib.pm_frame_ring_create(8, 6, 5000, ctypes.byref(h)) == 1 and not h.value
    assert lib.pm_frame_ring_put(NULL, 0, FAKE, NULL) == 1
    assert lib.pm_frame_ring_frame(NULL, 0, None, None) == 1
    n = ctypes.c_int64(0)
    assert lib.pm_frame_ring_puts(NULL, ctypes.byref(n)) == 1
    R = (ctypes.c_double * 9)(1, 0, 0, 0, 1, 0, 0, 0, 1)
    T = (ctypes.c_double * 3)(0, 0, 0)
    assert lib.pm_mapper_init_reference_frame(NULL, NULL, 0) == 1
    assert lib.pm_mapper_update_frame(NULL, NULL, 0, R, T, R, T) == 1
    assert lib.pm_mapper_output_fusion_cpu(NULL, FAKE, R, T, 1.0, 1.0, 0.0, 0.0) == 1
    assert C.last_error() != ""


def test_homography_against_explicit_projection():
    """h p + idep t must be proportional to the projection of the back-projected point
    (stereo_mapper.cpp:79-80 with camera->world poses)."""
    from probabilistic_mapping_amd.pipeline import homography
    rng = np.random.default_rng(0)
    K = np.array([[460.0, 0.3, 376.0], [0, 455.0, 240.0], [0, 0, 1]])
    from probabilistic_mapping_amd.synth import axis_angle
    Rl, Rr = axis_angle(rng.normal(size=3), 0.1), axis_angle(rng.normal(size=3), 0.2)
    Tl, Tr = rng.normal(size=3), rng.normal(size=3)
    h, t = homography(K, Rl, Tl, Rr, Tr)
    h = h.reshape(3, 3).astype(np.float64)
    for _ in range(20):
        p = np.array([rng.uniform(0, 752), rng.uniform(0, 480), 1.0])
        idep = rng.uniform(0.1, 2.0)
        Xl = np.linalg.inv(K) @ p / idep
        Xw = Rl @ Xl + Tl
        q = K @ (Rr.T @ (Xw - Tr))
        w = h @ p + idep * t
        np.testing.assert_allclose(w[:2] / w[2], q[:2] / q[2], rtol=2e-5)


def test_homography_rejects_bad_input():
    from probabilistic_mapping_amd.pipeline import homography
    with pytest.raises(C.PmError):
        homography(np.zeros((3, 3)), np.eye(3), np.zeros(3), np.eye(3), np.zeros(3))
    with pytest.raises(C.PmError):
        homography(np.eye(3), np.full((3, 3), np.nan), np.zeros(3), np.eye(3), np.zeros(3))
