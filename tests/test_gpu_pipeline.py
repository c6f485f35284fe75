"""End-to-end parity of the GPU pipeline against the FP64 oracle (needs a GPU).

Acceptance bars for the plain FP32 form:
  * WTA plane index identical on >= 99.9 % of pixels;
  * filtered depth category (rejected / non-overlap / value) identical on >= 99.9 % of pixels;
  * subpixel depth, fused mean, variance and inlier ratio: relative error <= 1e-4 on the
    pixels where both sides made the same discrete decisions, judged at the 99.9th
    percentile (the remaining 0.1 % sit on FP32-vs-FP64 decision boundaries of the
    parabola's curvature and are bounded separately).
"""
import json

import numpy as np
import pytest

from oracle.sequence import homography as np_homography
from tests import parity

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _report(name, stats):
    print(f"[parity] {name}: {json.dumps(stats)}", flush=True)


def _assert_depth(st):
    assert st["category_mismatch"] <= 1e-3
    assert st["rel"]["p999"] <= 1e-4


def test_stereo_mapper_small_scene(small_scene):
    """The drop-in StereoMapper on the 320x240, 2-source, 32-plane, 4-path workload."""
    from probabilistic_mapping_amd.mapper import StereoMapper, set_knobs
    import oracle
    s = small_scene
    K = s["scene"].K
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    old = set_knobs(idep_min=s["idep_min"], sgm_paths=4, median=0, **s["knobs"])
    try:
        m = StereoMapper(320, 240, D)
        m.initIntrinsic(K, 0.0, ds)
        m.initReference(s["ref"])
        for img, (R, T) in zip(s["srcs"], s["src_poses"]):
            m.update(img, *s["ref_pose"], R, T)
        g = m.output()
        g_cost = m.download("raw_cost")
        g_agg = m.download("sgm_cost")
    finally:
        set_knobs(**old)
    homs = [np_homography(K, *s["ref_pose"], R, T) for R, T in s["src_poses"]]
    o_cost = oracle.calc_cost(s["ref"], s["srcs"], homs, s["idep_min"], ds, D, "f64")
    o_agg = oracle.sgm(s["ref"], o_cost, s["knobs"], 4, "f64")
    o = oracle.filter_cost(o_agg, s["knobs"]["var_scale"], s["idep_min"], ds, "f64")
    wta_agree = float((parity.wta_index(g_agg) == parity.wta_index(o_agg)).mean())
    st = {"wta_agree": wta_agree, "cost": parity.cost_parity(g_cost, o_cost),
          "depth": parity.depth_parity(g, o)}
    _report("mapper small", st)
    assert wta_agree >= 0.999
    _assert_depth(st["depth"])


def test_mapper_host_images_reusable_after_call(small_scene):
    """initReference and update return before the device has used the image (the upload is
    asynchronous), so they must take their own copy first: overwriting the caller's buffer
    right after each call changes nothing.  Through the C ABI (pm_mapper_init_reference,
    pm_mapper_update), against a mapper fed buffers that are never touched."""
    from probabilistic_mapping_amd.mapper import StereoMapper, set_knobs
    s = small_scene
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    noise = np.random.default_rng(11)
    old = set_knobs(idep_min=s["idep_min"], sgm_paths=4, median=0, **s["knobs"])
    try:
        def run(clobber):
            m = StereoMapper(320, 240, D)
            m.initIntrinsic(s["scene"].K, 0.0, ds)
            buf = np.array(s["ref"], dtype=np.uint8, copy=True)
            m.initReference(buf)
            if clobber:
                buf[...] = noise.integers(0, 256, buf.shape, dtype=np.uint8)
            for img, (R, T) in zip(s["srcs"], s["src_poses"]):
                buf = np.array(img, dtype=np.uint8, copy=True)
                m.update(buf, *s["ref_pose"], R, T)
                if clobber:
                    buf[...] = noise.integers(0, 256, buf.shape, dtype=np.uint8)
            return m.output(), m.download("raw_cost")
        d0, c0 = run(False)
        d1, c1 = run(True)
    finally:
        set_knobs(**old)
    assert np.array_equal(c0, c1)
    assert np.array_equal(d0, d1, equal_nan=True)


def test_stereo_mapper_fusion_semantics(small_scene):
    """outputFusion: first keyframe returns zeros and only resets; later calls return the
    filtered (un-fused) depth, and the fused depth is a separate buffer."""
    from probabilistic_mapping_amd.mapper import StereoMapper, set_knobs
    s = small_scene
    K = s["scene"].K
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    old = set_knobs(idep_min=s["idep_min"], sgm_paths=4, median=0, **s["knobs"])
    try:
        m = StereoMapper(320, 240, D)
        m.initIntrinsic(K, 0.0, ds)
        fx, fy, cx, cy = s["scene"].intrinsics
        outs = []
        for k in range(3):
            m.initReference(s["ref"])
            for img, (R, T) in zip(s["srcs"], s["src_poses"]):
                m.update(img, *s["ref_pose"], R, T)
            outs.append(m.outputFusion(*s["ref_pose"], fx, fy, cx, cy))
        dep = m.download("dep")
        fused = m.download("fuse_depth")
        alpha = m.download("alpha")
    finally:
        set_knobs(**old)
    assert np.all(outs[0] == 0)
    assert np.array_equal(outs[1], dep) and np.array_equal(outs[2], dep)
    # identical keyframes at an identical pose: the third call fuses an initialised state
    assert (alpha >= 1e-5).mean() > 0.3
    assert not np.array_equal(fused, dep)


def test_case_modes(small_scene):
    """outputFusionCPU CASE 0 / 1 / 2 (stereo_mapper.cpp:628-719)."""
    from probabilistic_mapping_amd.mapper import StereoMapper, set_knobs
    import oracle
    s = small_scene
    K = s["scene"].K
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    fx, fy, cx, cy = s["scene"].intrinsics
    old = set_knobs(idep_min=s["idep_min"], sgm_paths=4, median=0, **s["knobs"])
    res = {}
    try:
        m = StereoMapper(320, 240, D)
        m.initIntrinsic(K, 0.0, ds)
        m.initReference(s["ref"])
        for img, (R, T) in zip(s["srcs"], s["src_poses"]):
            m.update(img, *s["ref_pose"], R, T)
        for mode in (0, 1, 2):
            set_knobs(case_mode=mode)
            res[mode] = m.outputFusionCPU(*s["ref_pose"], fx, fy, cx, cy)
        raw = m.download("raw_cost")
        agg = m.download("sgm_cost")
        dep = m.download("dep")
    finally:
        set_knobs(**old)

    def wta(c):
        idx = np.argmin(c, axis=2)
        mn = c.min(axis=2)
        return np.where(mn < 1e-5, 10000.0, np.where((idx == 0) | (idx == D - 1), 0.0,
                                                      1.0 / (s["idep_min"] + idx * ds)))
    assert np.allclose(res[0], wta(raw), rtol=1e-6, atol=0)
    assert np.allclose(res[1], wta(agg), rtol=1e-6, atol=0)
    assert np.array_equal(res[2], dep)


def test_mapper_call_order_errors():
    """update before initReference or initIntrinsic, output before initReference and a
    wrong-size image raise PM_ERR_INVALID_STATE / PM_ERR_INVALID_ARGUMENT instead of exiting
    (the reference would dereference uninitialised members); the mapper keeps working."""
    from probabilistic_mapping_amd import _capi as C
    from probabilistic_mapping_amd.mapper import StereoMapper
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (48, 64), dtype=np.uint8)
    I3, z = np.eye(3), np.zeros(3)
    m = StereoMapper(64, 48, 16)
    with pytest.raises(C.PmError) as e:
        m.output()
    assert e.value.status == C.PM_ERR_INVALID_STATE
    with pytest.raises(C.PmError) as e:
        m.update(img, I3, z, I3, z)
    assert e.value.status == C.PM_ERR_INVALID_STATE
    m.initReference(img)
    with pytest.raises(C.PmError) as e:
        m.update(img, I3, z, I3, z)                      # initIntrinsic not called
    assert e.value.status == C.PM_ERR_INVALID_STATE
    with pytest.raises(C.PmError) as e:
        m.initReference(np.zeros((48, 65), np.uint8))
    assert e.value.status == C.PM_ERR_INVALID_ARGUMENT
    K = np.array([[60.0, 0, 32], [0, 60.0, 24], [0, 0, 1]])
    m.initIntrinsic(K, 0.0, 1.0 / 15)
    m.update(img, I3, z, I3, z + np.array([0.05, 0, 0]))
    d = m.output()
    assert d.shape == (48, 64) and np.isfinite(d).all()
    m.close()


def test_case3_cpu_rules_sequence():
    """outputFusionCPU CASE 3 over 6 keyframes (stereo_mapper.cpp:719-800): the first call
    resets and returns zeros, each later one fuses the mapper's own filtered depth with the
    reference's CPU rules; result and state follow the oracle fed the same filtered depths
    and the double-precision pose composition, from the oracle's own state."""
    from probabilistic_mapping_amd import synth
    from probabilistic_mapping_amd.mapper import StereoMapper, set_knobs
    import oracle
    D = 32
    s = synth.trajectory_scene(7, width=160, height=120, n_frames=6, n_sources=2, planes=D,
                               fx=110.0)
    K = s["scene"].K
    fx, fy, cx, cy = (float(K[0, 0]), float(K[1, 1]), float(K[0, 2]), float(K[1, 2]))
    H, W = s["images"][0].shape
    old = set_knobs(idep_min=s["idep_min"], case_mode=3, **s["knobs"])
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    ost = [{k: np.zeros((H, W)) for k in ("alpha", "beta", "mu", "sig")} for _ in range(2)]
    pre = 0
    prev = None
    try:
        m = StereoMapper(W, H, D)
        m.initIntrinsic(K, 0.0, ds)
        for i in range(len(s["images"])):
            R, T = (np.asarray(a, np.float64) for a in s["poses"][i])
            T = T.reshape(3)
            m.initReference(s["images"][i])
            for j in s["sources"][i]:
                m.update(s["images"][j], R, T, *s["poses"][j])
            got = m.outputFusionCPU(R, T, fx, fy, cx, cy)
            if prev is None:
                assert not got.any()
                prev = (R, T)
                continue
            od = m.download("dep").astype(np.float64)
            r = R.T @ prev[0]
            t = R.T @ (prev[1] - T)
            otab = oracle.depth_fuse_cpu_rules(od, ost[pre], ost[pre ^ 1], r, t,
                                               (fx, fy, cx, cy), "f64")
            gstate = {k: m.download("cpu_" + k)
                      for k in ("alpha", "beta", "mu", "sig")}
            st = parity.state_parity(gstate, ost[pre ^ 1])
            st["table_mismatch"] = float((m.download("cpu_table") != otab).mean())
            st["fused"] = parity.depth_parity(got, od)
            _report(f"CASE 3 keyframe {i}", st)
            assert st["table_mismatch"] <= 1e-3
            assert st["init_mismatch"] <= 1e-3
            for k in ("mu", "sig", "inlier"):
                assert st[k]["p999"] <= 1e-4, k
            assert st["fused"]["category_mismatch"] <= 1e-3
            prev = (R, T)
            pre ^= 1
        m.close()
    finally:
        set_knobs(**old)


def test_small_scene_fusion_step_against_oracle(small_scene):
    """The 320x240, 2-source, 32-plane, 4-path workload through the drop-in StereoMapper with
    one Bayesian fusion step from a uniform prior, against the FP64 oracle.  With the
    reference's semantics the first outputFusion call only resets (stereo_mapper.cpp:128-146),
    the second initialises every measured pixel to alpha = beta = 10 (inlier prior 0.5),
    mu = d, sigma^2 = 2500 (depth_fusion.cu:161-174), and the third applies the update
    (:176-224); the keyframe is re-observed at the same pose each time."""
    from probabilistic_mapping_amd.mapper import StereoMapper, set_knobs
    import oracle
    s = small_scene
    K = s["scene"].K
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    fx, fy, cx, cy = s["scene"].intrinsics
    old = set_knobs(idep_min=s["idep_min"], sgm_paths=4, median=0, **s["knobs"])
    got = []
    try:
        m = StereoMapper(320, 240, D)
        m.initIntrinsic(K, 0.0, ds)
        for k in range(3):
            m.initReference(s["ref"])
            for img, (R, T) in zip(s["srcs"], s["src_poses"]):
                m.update(img, *s["ref_pose"], R, T)
            m.outputFusion(*s["ref_pose"], fx, fy, cx, cy)
            got.append({"fused": m.download("fuse_depth"), "dep": m.download("dep"),
                        "table": m.download("table"),
                        "state": {k2: m.download(k2) for k2 in ("alpha", "beta", "mu", "sig")}})
    finally:
        set_knobs(**old)
    homs = [np_homography(K, *s["ref_pose"], R, T) for R, T in s["src_poses"]]
    o_cost = oracle.calc_cost(s["ref"], s["srcs"], homs, s["idep_min"], ds, D, "f64")
    o_agg = oracle.sgm(s["ref"], o_cost, s["knobs"], 4, "f64")
    o_dep = oracle.filter_cost(o_agg, s["knobs"]["var_scale"], s["idep_min"], ds, "f64")
    H, W = o_dep.shape
    st = [{k: np.zeros((H, W)) for k in ("alpha", "beta", "mu", "sig")} for _ in range(2)]
    r, t = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    fused, tables = [], []
    for step in range(2):   # calls 2 and 3 (call 1 only resets)
        d = o_dep.copy()
        tables.append(oracle.depth_fuse(d, st[step], st[step ^ 1], r, t, (fx, fy, cx, cy),
                                        "f64"))
        fused.append((d, {k: v.copy() for k, v in st[step ^ 1].items()}))
    dep_st = parity.depth_parity(got[1]["dep"], o_dep)
    _report("small fusion: filtered depth", dep_st)
    _assert_depth(dep_st)
    for call, (od, ostate), otab in zip((1, 2), fused, tables):
        g = got[call]
        st_ = parity.state_parity(g["state"], ostate)
        st_["fused"] = parity.depth_parity(g["fused"], od)
        st_["table_mismatch"] = float((g["table"] != otab).mean())
        _report(f"small fusion call {call + 1}", st_)
        assert st_["init_mismatch"] <= 1e-3
        assert st_["table_mismatch"] <= 1e-3
        for k in ("mu", "sig", "inlier"):
            assert st_[k]["p999"] <= 1e-4, k
        assert st_["fused"]["category_mismatch"] <= 1e-3
        assert st_["fused"]["rel"]["p999"] <= 1e-4
    # the initialising call outputs the measurement; the updating call moves the inlier
    # ratio away from the uniform prior
    a, b = got[1]["state"]["alpha"], got[1]["state"]["beta"]
    init = a >= 1e-5
    assert np.all(a[init] == 10) and np.all(b[init] == 10)
    a2, b2 = got[2]["state"]["alpha"], got[2]["state"]["beta"]
    upd = (a2 >= 1e-5) & init
    assert np.abs(a2[upd] / (a2[upd] + b2[upd]) - 0.5).mean() > 1e-3


def test_mapper_trajectory_equals_sequence(trajectory_scene):
    """The C++ StereoMapper (host images in: initReference, update per source, outputFusion)
    over 6 keyframes of the rotating 752x480 trajectory gives the device-resident sequence's
    depth, fused depth and state bit for bit; the mapper composes R_p2c = R^T R_prev,
    t_p2c = R^T (t_prev - t) in C++ (stereo_mapper.cpp:155-156), the sequence in numpy."""
    from probabilistic_mapping_amd.mapper import StereoMapper, set_knobs
    from probabilistic_mapping_amd.pipeline import KeyframeSequence
    s = trajectory_scene
    n = 6
    K = s["scene"].K
    D = s["planes"]
    ds = (s["idep_max"] - s["idep_min"]) / (D - 1)
    fx, fy, cx, cy = s["scene"].intrinsics
    seq = KeyframeSequence(s["images"], s["poses"], s["sources"], K, D, s["idep_min"],
                           s["idep_max"], s["knobs"], s["paths"], cost_type="f32", median=True)
    seq.reset()
    old = set_knobs(idep_min=s["idep_min"], sgm_paths=s["paths"], median=1, **s["knobs"])
    try:
        H, W = s["images"][0].shape
        m = StereoMapper(W, H, D)
        m.initIntrinsic(K, 0.0, ds)
        for i in range(n):
            R, T = (np.asarray(a, np.float64) for a in s["poses"][i])
            m.initReference(s["images"][i])
            for j in s["sources"][i]:
                m.update(s["images"][j], R, T, *s["poses"][j])
            res = m.outputFusion(R, T.reshape(3), fx, fy, cx, cy)
            seq.run_keyframe(i)
            torch.cuda.synchronize()
            if i == 0:
                assert not res.any()
                continue
            assert np.array_equal(res, seq.dep.cpu().numpy()), i
            assert np.array_equal(m.download("fuse_depth"), seq.fused.cpu().numpy()), i
            for k in ("alpha", "beta", "mu", "sig"):
                assert np.array_equal(m.download(k), seq.state[k].cpu().numpy()), (i, k)
        m.close()
    finally:
        set_knobs(**old)
